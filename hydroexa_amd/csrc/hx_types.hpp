// hx_types.hpp -- types shared by the host C-ABI and both arithmetic builds of
// the kernels (exact: correctly rounded divisions and no FMA contraction,
// bit-identical to the FP64 CPU restatement; fast: shared Newton-refined
// reciprocals and contraction, held to 1e-12 relative by the parity tests).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <cstring>

namespace hx {

constexpr int kWaves = 4;              // wavefronts per workgroup
constexpr int kBlock = 64 * kWaves;
constexpr int kRoeCols = 62;           // columns updated per plain-Roe strip
constexpr int kHydroCols = 60;         // columns updated per hydrostatic strip

// Device view of one box (hx_patch with a typed element).
template <typename R>
struct PatchView {
    const R* h;
    const R* hu;
    const R* hv;
    const R* eta;   // may be null: eta = h + z
    const R* z;
    const R* n;
    int64_t pitch;
    int nx, ny;
    int bl0, bl1, bh0, bh1;   // per side: 0 reflect, 1 transmit, -1 stored ghosts
};

enum : unsigned { CFL_WET = 1u, CFL_DRY = 2u, CFL_NODATA = 4u };

// One CFL reduction, spread over kCflSubs partial slots, each on its own
// 128-byte line (device-scope atomics on one line serialize, whatever their
// addresses): a workgroup folds its partial into the sub-slot of the CU it
// runs on (slot_sub_here).  With a single slot, the device-scope atomics of
// workgroups that finish together all queue on one line (about 30 ns each);
// at 384^2 to 2048^2 that queue lengthened a launch by 4-10 us.  a and the
// flags sit in one 16-byte word, which a reader loads at once
// (slot_value_wave, one sub-slot a lane); the host folds with slot_value.
// A launch whose own next step reads the slot may fold into sub-slot 0 only
// (StepArgs::cfl_spread = 0): its readers then load one line instead of 64.
constexpr int kCflSubs = 64;
struct alignas(128) CflSub {
    unsigned long long amax;   // bits of a non-negative double
    unsigned int flags;
    unsigned int pad;
    unsigned long long vmax;   // bits of a non-negative double
    unsigned char line[104];
};
struct CflSlot {
    CflSub sub[kCflSubs];
};
static_assert(sizeof(CflSub) == 128 && sizeof(CflSlot) == 128 * kCflSubs, "CFL slot layout");

// A slot's reduced value: the maximum of a and of the velocity (as the bits
// of non-negative doubles, which order as the values do) and the OR of the
// flags over its sub-slots.
struct CflValue {
    unsigned long long amax, vmax;
    unsigned flags;
};
__host__ __device__ inline CflValue slot_value(const CflSlot* s)
{
    CflValue r{0ull, 0ull, 0u};
    for (int k = 0; k < kCflSubs; ++k) {
        const CflSub& q = s->sub[k];
        r.amax = q.amax > r.amax ? q.amax : r.amax;
        r.vmax = q.vmax > r.vmax ? q.vmax : r.vmax;
        r.flags |= q.flags;
    }
    return r;
}
__host__ __device__ inline void sub_clear(CflSlot* s, int k)
{
    s->sub[k].amax = 0ull;
    s->sub[k].flags = 0u;
    s->sub[k].pad = 0u;
    s->sub[k].vmax = 0ull;
}
__host__ __device__ inline void slot_clear(CflSlot* s)
{
    for (int k = 0; k < kCflSubs; ++k) sub_clear(s, k);
}

// dt estimate (before cfl) from a reduced slot: fl(1/max a) equals the
// reference's min over cells of fl(1/a) because correctly rounded division
// is monotone; dry cells contribute 1e6 and NODATA cells 1e30
// (src/solvers/Roe.cpp:42-50).
template <typename R>
__host__ __device__ inline R slot_dt_est(unsigned long long abits, unsigned flags);
template <typename R>
__host__ __device__ inline R slot_dt_est(const CflValue& v)
{
    return slot_dt_est<R>(v.amax, v.flags);
}
template <typename R>
__host__ __device__ inline R slot_dt_est(unsigned long long abits, unsigned flags)
{
    double a;
    memcpy(&a, &abits, sizeof(a));
    R est = (sizeof(R) == 8) ? R(1.7976931348623157e308) : R(3.40282347e38f);
    if (flags & CFL_WET) est = R(1) / R(a);
    if ((flags & CFL_DRY) && R(1.e6) < est) est = R(1.e6);
    if ((flags & CFL_NODATA) && R(1.e30) < est) est = R(1.e30);
    return est;
}

template <typename R>
struct StepArgs {
    PatchView<R> p;
    R* ho;
    R* huo;
    R* hvo;
    int64_t opitch;
    R dx, dy;
    R rdx, rdy;        // 1/dx, 1/dy
    int dx_pow2;       // 1/dx and 1/dy are exact (dx, dy powers of two)
    int rows;          // rows marched per wavefront
    int nsx, nstrips;  // strips across x, total strips
    double dt_host;    // used when rd == nullptr
    double cfl;
    CflSlot* rd;       // slot holding the CFL reduction of this step's input
    CflSlot* acc;      // slot receiving the CFL reduction of the output
    CflSlot* clr;      // slot cleared for the step after next
    double* dt_hist;
    double* vel_hist;
    int* err;
    int step;
    // forward hydrostatic step: when non-null, the state after the flux
    // update and before friction and the dry clamp is stored here too (the
    // adjoint's recomputation), three planes uno_plane apart at pitch opitch
    R* uno = nullptr;
    int64_t uno_plane = 0;
    // dt of this step in device memory (a multi-level rule computed it);
    // used when non-null, with err checked first
    const double* dt_dev = nullptr;
    // output cells [sk_i0, sk_i1) x [sk_j0, sk_j1) are left out of the CFL
    // reduction (a later pass reduces them after they change)
    int sk_i0 = 0, sk_j0 = 0, sk_i1 = 0, sk_j1 = 0;
    // banded tiling (advance_kernel): when band_wgs > 0 the grid is
    // 8 * band_wgs workgroups and XCD x marches rows [x ny / 8, (x + 1) ny / 8)
    // in strips of `rows` rows, the last tail_band of them in strips of
    // tail_rows; 0: uniform strips of `rows` rows
    int tail_rows = 0, tail_band = 0, band_wgs = 0;
    // when non-null, the effective fluxes of the patch's boundary faces as
    // its own cells take them: (2 ny + 2 nx) x 3 values, x-low faces
    // (i = 0) by row, x-high (i = nx), y-low (j = 0) by column, y-high
    // (j = ny); plain Roe D- on the low faces and D+ on the high faces,
    // hydrostatic 0.5 FL + 0.5 FR
    R* ef = nullptr;
    // 1: the CFL partials fold into the sub-slot of the CU (acc) and every
    // sub-slot of rd is read; 0: sub-slot 0 only, both ways
    int cfl_spread = 1;
};

// The strip of the banded tiling that item i (workgroup index within its
// XCD times kWaves plus the wavefront) of XCD `xcd` marches: its column
// index sxi and rows [y0, y1); none when y0 >= y1.  XCD x takes the rows
// [x ny / 8, (x + 1) ny / 8) bottom to top: strips of a.rows rows, then, over
// at least the band's last a.tail_band rows, strips of a.tail_rows rows.
template <typename R>
__host__ __device__ inline void band_strip(const StepArgs<R>& a, unsigned xcd, int i, int& sxi,
                                           int& y0, int& y1)
{
    const int ny = a.p.ny;
    const int yb0 = (int)(((int64_t)xcd * ny) >> 3);
    const int yb1 = (int)(((int64_t)(xcd + 1) * ny) >> 3);
    const int nm0 = (yb1 - yb0 - a.tail_band) / a.rows;
    const int nm = nm0 > 0 ? nm0 : 0;
    const int sr = i / a.nsx;
    sxi = i - sr * a.nsx;
    if (sr < nm) {
        y0 = yb0 + sr * a.rows;
        y1 = y0 + a.rows;
    } else {
        y0 = yb0 + nm * a.rows + (sr - nm) * a.tail_rows;
        y1 = (y0 + a.tail_rows < yb1) ? y0 + a.tail_rows : yb1;
    }
}

// Faces whose effective fluxes build a flux register: a device list of
// {dir, i, j, side}, or (faces == null) the faces around the fine box of
// coarse cells [ci0, ci0 + ncx) x [cj0, cj0 + ncy) -- box_fine 0: the
// coarse faces x-low, x-high, y-low, y-high with the coarse side's view;
// box_fine 1: the two fine faces of each, in the fine box's own indices,
// with the fine side's view.  Scales per direction, times *dt when dt is
// non-null.
struct FaceReq {
    const int* faces;
    int nfaces, pair, accumulate;
    int box_fine, ci0, cj0, ncx, ncy;
    double sx, sy;
    const double* dt;
    const int* err;     // non-null and set: nothing is done (a stopped run)
};

// One box of an AMR level in the level's patch table: every per-box
// operation of a substep (the advance, the CFL reduction, the ghost fill,
// the flux-register faces, reflux and average_down) runs as one launch over
// the table, each box owning a run of strips or blocks from its first index
// below.
struct AmrPatch {
    int i0, j0, w, h;     // box in the level's cells
    int rows, nsx;        // advance strips of the box
    int strip0;           // first strip of the box in the launch
    int cfl0, cflx;       // CFL reduction: first block, blocks across x
    int fill0;            // ghost fill: first block
    int face0;            // flux-register rows (and reflux faces): first block
    int avg0;             // average_down: first block
    int reg;              // offset of the box's flux register (elements)
};

// A level's dense arrays as the batched kernels see them: each buffer holds
// the h, hu, hv planes plane elements apart, the terrain its z and roughness
// planes; a cell (i, j) lies at (j + 2) pitch + (i + 2) (two padding cells),
// pitch = nx + 4.  Box sides on the domain boundary take the physical BCs.
template <typename R>
struct AmrLevelArgs {
    const R* u_in;        // the state read
    R* u_out;             // the state written (the advance)
    const R* terrain;
    const uint8_t* mask;  // 1 on the cells the level owns (pitch as the arrays)
    int64_t pitch, plane;
    int nx, ny;
    int bclo0, bclo1, bchi0, bchi1;
    R dx, dy, rdx, rdy;
    int dx_pow2;
};

// View of box (i0, j0, w, h) of a level on state u.
template <typename R>
__host__ __device__ inline PatchView<R> amr_view(const AmrLevelArgs<R>& L, const R* u, int i0,
                                                 int j0, int w, int h)
{
    const int64_t off = (int64_t)(j0 + 2) * L.pitch + (i0 + 2);
    PatchView<R> p;
    p.h = u + off;
    p.hu = u + L.plane + off;
    p.hv = u + 2 * L.plane + off;
    p.eta = nullptr;
    p.z = L.terrain + off;
    p.n = L.terrain + L.plane + off;
    p.pitch = L.pitch;
    p.nx = w;
    p.ny = h;
    p.bl0 = (i0 == 0) ? L.bclo0 : -1;
    p.bl1 = (j0 == 0) ? L.bclo1 : -1;
    p.bh0 = (i0 + w == L.nx) ? L.bchi0 : -1;
    p.bh1 = (j0 + h == L.ny) ? L.bchi1 : -1;
    return p;
}

// The box of a patch table that owns index key: the last k with
// first(t[k]) <= key (first indices ascend).  key is uniform across the
// caller's wavefront or workgroup, so the search is scalar.
template <typename F>
__device__ inline int amr_find(const AmrPatch* t, int n, int key, F first)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first(t[mid]) <= key) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Scenario table of a batched adjoint rollout (FP64): one entry per
// scenario, with the buffers its forward and its segment recomputations
// step through.  State k of the forward lives in ckpt[k / every] when k is a
// multiple of every, else in seg[k & 1]; inside the recomputation of the
// segment starting at s0, state t > s0 lives in seg[t - s0 - 1] and the
// hydrostatic step t stores its pre-friction state in un[t - s0].  Each
// slot is 3 planes of N doubles.
struct AdjScen {
    PatchView<double> p;       // geometry, BCs and terrain (h/hu/hv unused)
    double dx, dy, rdx, rdy;
    int dx_pow2;
    int rows, nsx, nstrips, strip0;     // the advance's strips
    int ntx, ntiles, tile0;             // the adjoint's tiles
    int ngauge, gauge0, nu, u0;         // gauges and distinct gauge cells
    int64_t N;
    double* ckpt;
    double* seg;
    double* un;
    double* lam;               // 2 slots: adjoint ping-pong
    double* lun;               // 1 slot (hydrostatic)
    double* out_h;             // the forward's final state (caller's planes)
    double* out_hu;
    double* out_hv;
    double* series;            // the forward's gauge series (caller's), [step][gauge]
    const double* in_h;        // the forward's initial state (caller's planes)
    const double* in_hu;
    const double* in_hv;
    const double* gs;          // the backward's series adjoint (caller's, optional)
    const double* gfin;        // the backward's final-state adjoint (caller's, optional)
    double* gstate;            // the backward's gradients (caller's): initial state,
    double* gz;                //   bed and roughness
    double* gn;
    double* dts;               // [nsteps] dt of every step
    double* vel;               // [nsteps] max|vel| before every step (CFL mode)
    CflSlot* slots;            // 3-slot CFL ring
    int* err;                  // dt collapse: step + 1
};

// One launch over a batch: step t of the forward (s0 < 0) or of the
// recomputation of the segment that starts at s0.
struct AdjBatch {
    const AdjScen* sc;
    const int* strip_start;    // [nscen + 1]
    int nscen, total;
    int nsteps, every;
    int t, s0;
    int use_cfl;               // forward only: dt from the CFL ring, else dts[t]
    int ring;                  // t modulo 3
    double cfl;
};

// Occupancy variants compiled for the advance kernel (minimum waves per
// SIMD requested from the register allocator).
constexpr int kOccVariants = 3;
constexpr int kOcc[kOccVariants] = {1, 3, 4};

// Host launchers of the two arithmetic builds.
namespace exact {
template <typename R>
int launch_advance(int solver, int occ, const StepArgs<R>& a, hipStream_t st);
template <typename R>
int simd_waves(int solver, int occ);
template <typename R>
int launch_bflux(int solver, const PatchView<R>& p, R dx, R dy, const FaceReq& q, R* out,
                 hipStream_t st);
}
namespace fast {
template <typename R>
int launch_advance(int solver, int occ, const StepArgs<R>& a, hipStream_t st);
template <typename R>
int simd_waves(int solver, int occ);
template <typename R>
int launch_bflux(int solver, const PatchView<R>& p, R dx, R dy, const FaceReq& q, R* out,
                 hipStream_t st);
}

}  // namespace hx

namespace hx {

// One scenario of a batched ensemble (device-resident descriptor table).
// Buffers A and B alternate as input and output: even steps read A.
template <typename R>
struct ScenDesc {
    PatchView<R> p;        // geometry, BCs, terrain, and buffer A in h/hu/hv
    R* bh;                 // buffer B
    R* bhu;
    R* bhv;
    R dx, dy, rdx, rdy;
    int dx_pow2;
    int nsx;               // strips across x
    int strip0;            // first global strip index of this scenario
    int nstrips;
    CflSlot* slots;        // 3-slot ring of this scenario
    int* err;              // dt-collapse flag (step + 1), 0 if none
    double* dt_hist;       // [hist_cap]
    double* vel_hist;      // [hist_cap]
};

template <typename R>
struct EnsArgs {
    const ScenDesc<R>* desc;
    const int* strip_start;   // [nscen + 1] prefix of nstrips
    int nscen;
    int total;                // total strips
    int rows;
    int parity;               // 0: A -> B, 1: B -> A
    int ring;                 // step index modulo 3 (slot ring position)
    const int* step_base;     // device step counter (graph cycles advance it)
    int step_off;             // step = *step_base + step_off
    double cfl;
    unsigned* queue;          // 3-counter ring of the persistent work queue
    int persistent;           // 0: one wavefront per strip; 1: persistent waves pull strips
};

namespace exact {
template <typename R>
int launch_ensemble(int solver, int occ, const EnsArgs<R>& e, int grid, hipStream_t st);
int launch_adj_batch(int solver, int occ, const AdjBatch& b, hipStream_t st);
template <typename R>
int launch_amr_advance(int solver, int occ, const AmrLevelArgs<R>& L, const AmrPatch* tab,
                       int npatch, int strips, R dt, hipStream_t st);
template <typename R>
int launch_amr_faces(int solver, const AmrLevelArgs<R>& L, const AmrPatch* tab, int npatch,
                     int blocks, int fine_box, double sx, double sy, R* regs, hipStream_t st);
}
namespace fast {
template <typename R>
int launch_ensemble(int solver, int occ, const EnsArgs<R>& e, int grid, hipStream_t st);
int launch_adj_batch(int solver, int occ, const AdjBatch& b, hipStream_t st);
template <typename R>
int launch_amr_advance(int solver, int occ, const AmrLevelArgs<R>& L, const AmrPatch* tab,
                       int npatch, int strips, R dt, hipStream_t st);
template <typename R>
int launch_amr_faces(int solver, const AmrLevelArgs<R>& L, const AmrPatch* tab, int npatch,
                     int blocks, int fine_box, double sx, double sy, R* regs, hipStream_t st);
}

}  // namespace hx
