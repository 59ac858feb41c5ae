// hx_kernels.hip -- exact arithmetic build of the fused advance, the
// auxiliary kernels (CFL reduction, tagging, masked_average_down, pruning)
// and the C-ABI declared in include/hydroexa.h.  Compiled with
// -ffp-contract=off; the fast build of the advance lives in hx_fast.hip.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/hydroexa.h"
#include "hx_types.hpp"
#include "hx_adjoint.hpp"
#include "hx_host.hpp"

namespace hx {
#define HX_NS exact
#define HX_FAST 0
#include "hx_device.hpp"
#include "hx_advance.inc"
#undef HX_NS
#undef HX_FAST

using namespace exact;

// CFL reduction of a stored state (compute_dt_Impl's ReduceOps pass).  A
// thread owns one column of a run of `rows` rows (kCflRows unless the launch
// shortens them, blockIdx.y); the loads of kCflUnroll rows are issued before
// any is used.
constexpr int kCflRows = 16;
constexpr int kCflUnroll = 4;

template <typename R>
__device__ __forceinline__ void cfl_block(const PatchView<R>& p, R dx, R dy, R rdx, R rdy, int mul,
                                          CflSlot* slot, int bx, int by, bool spread = true,
                                          int rows = kCflRows)
{
    CflAcc<R> acc;
    acc.init();
    const int i = bx * kBlock + threadIdx.x;
    const int j0 = by * rows;
    const int j1 = min(j0 + rows, p.ny);
    if (i < p.nx) {
        int j = j0;
        for (; j + kCflUnroll <= j1; j += kCflUnroll) {
            R h[kCflUnroll], qx[kCflUnroll], qy[kCflUnroll], z[kCflUnroll];
#pragma unroll
            for (int u = 0; u < kCflUnroll; ++u) {
                const int64_t s = (int64_t)(j + u) * p.pitch + i;
                h[u] = p.h[s]; qx[u] = p.hu[s]; qy[u] = p.hv[s]; z[u] = p.z[s];
            }
#pragma unroll
            for (int u = 0; u < kCflUnroll; ++u)
                acc.add(h[u], qx[u], qy[u], z[u], dx, dy, rdx, rdy, mul != 0);
        }
        for (; j < j1; ++j) {
            const int64_t s = (int64_t)j * p.pitch + i;
            acc.add(p.h[s], p.hu[s], p.hv[s], p.z[s], dx, dy, rdx, rdy, mul != 0);
        }
    }
    cfl_commit(acc, slot, spread);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void cfl_kernel(PatchView<R> p, R dx, R dy, R rdx, R rdy,
                                                     int mul, CflSlot* slot, int spread, int rows)
{
    cfl_block(p, dx, dy, rdx, rdy, mul, slot, blockIdx.x, blockIdx.y, spread != 0, rows);
}

__global__ void clear_slots_kernel(CflSlot* s, int n, int* err)
{
    for (int q = threadIdx.x; q < n * kCflSubs; q += blockDim.x) sub_clear(&s[q / kCflSubs], q % kCflSubs);
    if (threadIdx.x == 0 && err) *err = 0;
}

// ---------------------------------------------------------------------------
// Tagging (src/solvers/Roe.cpp:121-179, HydrostaticRoe.cpp:109-169).  The
// halo is the state's own BC image: h with the scalar rule, momenta with the
// vector rules, terrain foextrap.
// ---------------------------------------------------------------------------
template <typename R>
struct TagCell { R h, qx, qy, z; };

template <typename R>
__device__ __forceinline__ TagCell<R> load_tag(const PatchView<R>& p, int i, int j)
{
    int si, ei, sj, ej;
    bool fx, fy;
    map_axis(i, p.nx, p.bl0, p.bh0, si, ei, fx);
    map_axis(j, p.ny, p.bl1, p.bh1, sj, ej, fy);
    const int64_t s = (int64_t)sj * p.pitch + si;
    const int64_t t = (int64_t)ej * p.pitch + ei;
    TagCell<R> c;
    c.h = p.h[s];
    const R qx = p.hu[s], qy = p.hv[s];
    c.qx = fx ? -qx : qx;
    c.qy = fy ? -qy : qy;
    c.z = p.z[t];
    return c;
}

template <typename R>
__device__ __forceinline__ R umag(const TagCell<R>& c)
{
    if (c.h < R(1.0e-6)) return R(0);
    const R u = c.qx / c.h, v = c.qy / c.h;
    return sqrt(u * u + v * v);
}

template <typename R>
__global__ __launch_bounds__(256) void tag_kernel(PatchView<R> p, int solver, R dx, R dy,
                                                  double time, R hth, R uth, R zth,
                                                  uint8_t* tags)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= p.nx || j >= p.ny) return;
    const TagCell<R> E = load_tag(p, i + 1, j), W = load_tag(p, i - 1, j);
    const TagCell<R> N = load_tag(p, i, j + 1), S = load_tag(p, i, j - 1);
    bool skip;
    if (solver == HX_SOLVER_ROE)
        skip = (E.z == R(-9999.0) || W.z == R(-9999.0) || N.z == R(-9999.0) || S.z == R(-9999.0));
    else
        skip = (fabs(E.z + R(9999.0)) < Kc<R>::tol10 || fabs(W.z + R(9999.0)) < Kc<R>::tol10 ||
                fabs(N.z + R(9999.0)) < Kc<R>::tol10 || fabs(S.z + R(9999.0)) < Kc<R>::tol10);
    uint8_t t = 0;
    if (!skip) {
        const R dhdx = (E.h - W.h) / (R(2) * dx);
        const R dhdy = (N.h - S.h) / (R(2) * dy);
        bool tg = sqrt(dhdx * dhdx + dhdy * dhdy) > hth;
        const bool early = (time <= 1.0e-8);
        if (solver == HX_SOLVER_ROE && early) {
            const R dzdx = (E.z - W.z) / (R(2) * dx), dzdy = (N.z - S.z) / (R(2) * dy);
            if (sqrt(dzdx * dzdx + dzdy * dzdy) > zth) tg = true;
        }
        if (!tg) {
            const R dudx = (umag(E) - umag(W)) / (R(2) * dx);
            const R dudy = (umag(N) - umag(S)) / (R(2) * dy);
            if (sqrt(dudx * dudx + dudy * dudy) > uth) tg = true;
            if (solver != HX_SOLVER_ROE && early) {
                const R dzdx = (E.z - W.z) / (R(2) * dx), dzdy = (N.z - S.z) / (R(2) * dy);
                if (sqrt(dzdx * dzdx + dzdy * dzdy) > zth) tg = true;
            }
        }
        t = tg ? 1 : 0;
    }
    tags[(int64_t)j * p.nx + i] = t;
}

// ---------------------------------------------------------------------------
// masked_average_down (include/utils/MaskedAverageDown.H:63-104): one thread
// per coarse cell and component, children visited jref-major as in the
// reference so the sums round identically.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void avgdown_kernel(const R* fine, int nxf, int nyf, int ncomp,
                                                      int rx, int ry, R fv, R nodata, R* crse,
                                                      int32_t* nvalid)
{
    const int nxc = nxf / rx, nyc = nyf / ry;
    const int64_t ncell = (int64_t)nxc * nyc;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ncell * ncomp) return;
    const int comp = (int)(q / ncell);
    const int64_t cc = q - (int64_t)comp * ncell;
    const int j = (int)(cc / nxc), i = (int)(cc - (int64_t)j * nxc);
    const R* f = fine + (int64_t)comp * nxf * nyf;
    R cd = R(0), cv = R(0);
    int nv = 0;
    for (int jr = 0; jr < ry; ++jr)
        for (int ir = 0; ir < rx; ++ir) {
            const R val = f[(int64_t)(j * ry + jr) * nxf + (i * rx + ir)];
            if (!isnan(val) && fabs(val - nodata) > R(1e-3)) {
                ++nv;
                cv += fv;
                cd += val * fv;
            }
        }
    crse[q] = (nv * 2 >= rx * ry && cv > R(0)) ? cd / cv : nodata;
    if (nvalid) nvalid[q] = nv;
}

// ---------------------------------------------------------------------------
// Pruning predicate (src/state/AmrMeshState_AMR.cpp:342-364): one workgroup
// per box scans its refined footprint; flags[2b] = has valid, [2b+1] = has
// NODATA (z <= -9998).
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void prune_kernel(const int32_t* boxes, const R* z, int nxs,
                                                    int nys, int sc, uint32_t* flags)
{
    const int b = blockIdx.x;
    int ilo = boxes[4 * b] * sc, jlo = boxes[4 * b + 1] * sc;
    int ihi = (boxes[4 * b + 2] + 1) * sc - 1, jhi = (boxes[4 * b + 3] + 1) * sc - 1;
    ilo = max(ilo, 0); jlo = max(jlo, 0);
    ihi = min(ihi, nxs - 1); jhi = min(jhi, nys - 1);
    const int w = ihi - ilo + 1, hgt = jhi - jlo + 1;
    uint32_t hv = 0, hn = 0;
    if (w > 0 && hgt > 0) {
        const int64_t n = (int64_t)w * hgt;
        for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
            const int jj = jlo + (int)(q / w), ii = ilo + (int)(q % w);
            const bool nd = z[(int64_t)jj * nxs + ii] <= R(-9998.0);
            if (nd) hn = 1; else hv = 1;
        }
    }
    hv = __any(hv) ? 1u : 0u;
    hn = __any(hn) ? 1u : 0u;
    if ((threadIdx.x & 63) == 0) {
        if (hv) atomicOr(&flags[2 * b], 1u);
        if (hn) atomicOr(&flags[2 * b + 1], 1u);
    }
}

__global__ void zero_u32_kernel(uint32_t* p, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0u;
}

// Ensemble set-up: clear every scenario's CFL ring and error flag, the work
// queue ring and the step counter (kernels, not memsets, so the same code is
// safe inside graph capture).
__global__ void ens_reset_kernel(CflSlot* slots, int nslots, int* err, int nscen,
                                 unsigned* queue, int* step_base)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nslots) slot_clear(slots + q);
    if (q < nscen) err[q] = 0;
    if (q < 3) queue[q] = 0u;
    if (q == 0) *step_base = 0;
}

__global__ void ens_tick_kernel(int* step_base, int inc)
{
    if (threadIdx.x == 0) *step_base += inc;
}

// Initial CFL reduction of every scenario (blockIdx.y = scenario) into
// slot 0 of its ring, read from buffer A.
template <typename R>
__global__ __launch_bounds__(kBlock) void ens_cfl_kernel(const ScenDesc<R>* desc)
{
    const ScenDesc<R>& d = desc[blockIdx.y];
    const PatchView<R>& p = d.p;
    CflAcc<R> acc;
    acc.init();
    const int64_t ncell = (int64_t)p.nx * p.ny;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ncell;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(q / p.nx), i = (int)(q - (int64_t)j * p.nx);
        const int64_t s = (int64_t)j * p.pitch + i;
        acc.add(p.h[s], p.hu[s], p.hv[s], p.z[s], d.dx, d.dy, d.rdx, d.rdy, d.dx_pow2 != 0);
    }
    cfl_commit(acc, d.slots);
}

// ---------------------------------------------------------------------------
// Conservation diagnostics (src/state/AmrMeshState_Utils.cpp:213-343): one
// fused pass over the valid cells of a level gives sum h dA, sum hu dA,
// sum hv dA, sum 0.5 (hu^2 + hv^2)/h dA and min h, over cells whose
// roughness is not NODATA and h > 0.  Sums are compensated (TwoSum) per
// thread, folded by a fixed tree per block and by one thread over the blocks,
// so the result depends only on the level size, not on scheduling.
// ---------------------------------------------------------------------------
constexpr int kConsBlocks = 1024;

struct KSum {
    double s, c;
};

__device__ __forceinline__ void kadd(KSum& a, double x)
{
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
    else a.c += (x - t) + a.s;
    a.s = t;
}

__device__ __forceinline__ void kmerge(KSum& a, const KSum& b)
{
    kadd(a, b.s);
    a.c += b.c;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void cons_kernel(PatchView<R> p, R dA, double* part)
{
    __shared__ KSum sh[4][kBlock];
    __shared__ double smin_[kBlock];
    KSum acc[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    double mn = 1.7976931348623157e308;
    const int64_t ncell = (int64_t)p.nx * p.ny;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < ncell;
         q += (int64_t)gridDim.x * kBlock) {
        const int j = (int)(q / p.nx), i = (int)(q - (int64_t)j * p.nx);
        const int64_t s = (int64_t)j * p.pitch + i;
        const R h = p.h[s], qx = p.hu[s], qy = p.hv[s];
        if (!(p.n[s] == R(-9999.0)) && h > R(0)) {
            kadd(acc[0], (double)(h * dA));
            kadd(acc[1], (double)(qx * dA));
            kadd(acc[2], (double)(qy * dA));
            kadd(acc[3], (double)(R(0.5) * (qx * qx + qy * qy) / h * dA));
            mn = ((double)h < mn) ? (double)h : mn;
        }
    }
    const int t = threadIdx.x;
    for (int k = 0; k < 4; ++k) sh[k][t] = acc[k];
    smin_[t] = mn;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int k = 0; k < 4; ++k) kmerge(sh[k][t], sh[k][t + w]);
            smin_[t] = (smin_[t + w] < smin_[t]) ? smin_[t + w] : smin_[t];
        }
        __syncthreads();
    }
    if (t == 0) {
        double* o = part + (size_t)blockIdx.x * 9;
        for (int k = 0; k < 4; ++k) { o[2 * k] = sh[k][0].s; o[2 * k + 1] = sh[k][0].c; }
        o[8] = smin_[0];
    }
}

// One workgroup of kConsBlocks threads folds the block partials by a fixed
// tree (thread t holds partial t), so the result is repeatable.
__global__ __launch_bounds__(kConsBlocks) void cons_final_kernel(const double* part, int nblk,
                                                                 double* out)
{
    __shared__ KSum sh[4][kConsBlocks];
    __shared__ double smin_[kConsBlocks];
    const int t = threadIdx.x;
    for (int k = 0; k < 4; ++k)
        sh[k][t] = (t < nblk) ? KSum{part[(size_t)t * 9 + 2 * k], part[(size_t)t * 9 + 2 * k + 1]}
                              : KSum{0.0, 0.0};
    smin_[t] = (t < nblk) ? part[(size_t)t * 9 + 8] : 1.7976931348623157e308;
    __syncthreads();
    for (int w = kConsBlocks / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int k = 0; k < 4; ++k) kmerge(sh[k][t], sh[k][t + w]);
            smin_[t] = (smin_[t + w] < smin_[t]) ? smin_[t + w] : smin_[t];
        }
        __syncthreads();
    }
    if (t == 0) {
        for (int k = 0; k < 4; ++k) out[k] = sh[k][0].s + sh[k][0].c;
        out[4] = smin_[0];
    }
}

// ---------------------------------------------------------------------------
// Two-level coupling for one fine box at refinement ratio 2 (the reference's
// AMR level coupling: src/state/AmrMeshState_AMR.cpp:667-844,
// src/state/AmrMeshState_Physics.cpp:50-147, and AMReX's FillPatchTwoLevels,
// FluxRegister and average_down that it calls).  The fine box covers coarse
// cells [ci0, ci0 + ncx) x [cj0, cj0 + ncy); its state and terrain carry
// 2 stored ghost layers.
// ---------------------------------------------------------------------------

// Coarse sample for the fine ghost fill: eta = h + z, and the momenta, at a
// coarse index that may lie outside the level (physical-BC source cell,
// normal momentum negated across a reflective side).
template <typename R>
__device__ __forceinline__ void coarse_q(const PatchView<R>& c, int i, int j, R& eta, R& qx, R& qy)
{
    int si, ei, sj, ej;
    bool fx, fy;
    map_axis(i, c.nx, c.bl0, c.bh0, si, ei, fx);
    map_axis(j, c.ny, c.bl1, c.bh1, sj, ej, fy);
    const int64_t s = (int64_t)sj * c.pitch + si;
    eta = c.h[s] + c.z[s];
    const R u = c.hu[s], v = c.hv[s];
    qx = fx ? -u : u;
    qy = fy ? -v : v;
}

template <typename R>
__device__ __forceinline__ void coarse_q_t(const PatchView<R>& o, const PatchView<R>& n, R alpha,
                                           int i, int j, R* q)
{
    R a[3], b[3];
    coarse_q(o, i, j, a[0], a[1], a[2]);
    if (alpha == R(0)) { q[0] = a[0]; q[1] = a[1]; q[2] = a[2]; return; }
    coarse_q(n, i, j, b[0], b[1], b[2]);
    if (alpha == R(1)) { q[0] = b[0]; q[1] = b[1]; q[2] = b[2]; return; }
    for (int k = 0; k < 3; ++k) q[k] = (R(1) - alpha) * a[k] + alpha * b[k];
}

// Monotonized-central limited slope of a coarse cell.
template <typename R>
__device__ __forceinline__ R mc_slope(R m, R c, R p)
{
    const R cen = R(0.5) * (p - m), fw = p - c, bw = c - m;
    const R lim = (fw * bw > R(0)) ? R(2) * fmin(fabs(fw), fabs(bw)) : R(0);
    return copysign(fmin(fabs(cen), lim), cen);
}

// One thread per fine ghost cell (2 layers around the box, corners too):
// conservative linear interpolation of eta, hu, hv from the time-blended
// coarse parent, then h = max(0, eta - z) on the fine ghost's own bed.
template <typename R>
__device__ __forceinline__ void fillfine_cell(const PatchView<R>& co, const PatchView<R>& cn,
                                              R alpha, int ci0, int cj0, int nxf, int nyf,
                                              int64_t fpitch, int whole, const R* __restrict__ zf,
                                              R* hf, R* huf, R* hvf,
                                              const uint8_t* __restrict__ skip, int64_t spitch,
                                              int t)
{
    const int W = nxf + 4, H = nyf + 4;
    const int n = whole ? W * H : W * H - nxf * nyf;
    if (t >= n) return;
    // whole: every cell of the box and its ghosts.  Otherwise the ghost ring:
    // rows -2, -1 and ny, ny+1 in full, then the side columns
    int fi, fj;
    if (whole) { fj = t / W - 2; fi = t % W - 2; }
    else if (t < 2 * W) { fj = t / W - 2; fi = t % W - 2; }
    else if (t < 4 * W) { const int u = t - 2 * W; fj = nyf + u / W; fi = u % W - 2; }
    else { const int u = t - 4 * W; fj = u / 4; const int k = u % 4; fi = (k < 2) ? k - 2 : nxf + k - 2; }
    // cells the fine level already holds (another box's valid cells) keep them
    if (skip && skip[(int64_t)fj * spitch + fi]) return;
    const int pi = ci0 + (fi >= 0 ? fi / 2 : -((1 - fi) / 2));
    const int pj = cj0 + (fj >= 0 ? fj / 2 : -((1 - fj) / 2));
    const R ox = ((fi & 1) == 0) ? R(-0.25) : R(0.25);
    const R oy = ((fj & 1) == 0) ? R(-0.25) : R(0.25);
    R qc[3], qw[3], qe[3], qs[3], qn[3];
    coarse_q_t(co, cn, alpha, pi, pj, qc);
    coarse_q_t(co, cn, alpha, pi - 1, pj, qw);
    coarse_q_t(co, cn, alpha, pi + 1, pj, qe);
    coarse_q_t(co, cn, alpha, pi, pj - 1, qs);
    coarse_q_t(co, cn, alpha, pi, pj + 1, qn);
    R v[3];
    for (int k = 0; k < 3; ++k)
        v[k] = qc[k] + mc_slope(qw[k], qc[k], qe[k]) * ox + mc_slope(qs[k], qc[k], qn[k]) * oy;
    const int64_t o = (int64_t)fj * fpitch + fi;
    const R x = v[0] - zf[o];
    hf[o] = (R(0) < x) ? x : R(0);
    huf[o] = v[1];
    hvf[o] = v[2];
}

template <typename R>
__global__ __launch_bounds__(kBlock) void fillfine_kernel(PatchView<R> co, PatchView<R> cn,
                                                          R alpha, int ci0, int cj0, int nxf,
                                                          int nyf, int64_t fpitch, int whole,
                                                          const R* __restrict__ zf, R* hf, R* huf,
                                                          R* hvf, const int* err,
                                                          const uint8_t* __restrict__ skip,
                                                          int64_t spitch)
{
    if (err && *err) return;
    fillfine_cell(co, cn, alpha, ci0, cj0, nxf, nyf, fpitch, whole, zf, hf, huf, hvf, skip, spitch,
                  (int)(blockIdx.x * kBlock + threadIdx.x));
}

// Reflux: register q of coarse face q holds -dt dA F_coarse + sum over the
// fine substeps and fine faces of dt_f dA_f F_fine.  The coarse cell outside
// the box beside the face gains (-reg)/vol on the box's low sides and
// reg/vol on its high sides, vol = dx dy (AMReX FluxRegister::Reflux).
// orient (-1: all) restricts to one side, numbered as AMReX's Orientation
// (0 x-low, 1 y-low, 2 x-high, 3 y-high), so several boxes can be refluxed
// side by side in AMReX's order; cells outside the level (a box side on the
// domain boundary) are left alone.
template <typename R>
__device__ __forceinline__ void reflux_face(int ci0, int cj0, int ncx, int ncy, int64_t pitch, R vol,
                                            const R* __restrict__ reg, R* h, R* hu, R* hv,
                                            int orient, int cnx, int cny, int q)
{
    const int nq = 2 * ncy + 2 * ncx;
    if (q >= nq) return;
    int i, j, side;
    R sgn;
    if (q < ncy) { i = ci0 - 1; j = cj0 + q; sgn = R(-1); side = 0; }
    else if (q < 2 * ncy) { i = ci0 + ncx; j = cj0 + q - ncy; sgn = R(1); side = 2; }
    else if (q < 2 * ncy + ncx) { i = ci0 + q - 2 * ncy; j = cj0 - 1; sgn = R(-1); side = 1; }
    else { i = ci0 + q - 2 * ncy - ncx; j = cj0 + ncy; sgn = R(1); side = 3; }
    if ((orient >= 0 && side != orient) || i < 0 || j < 0 || i >= cnx || j >= cny) return;
    const int64_t o = (int64_t)j * pitch + i;
    const R* r = reg + 3 * q;
    h[o] += (sgn * r[0]) / vol;
    hu[o] += (sgn * r[1]) / vol;
    hv[o] += (sgn * r[2]) / vol;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void reflux_kernel(int ci0, int cj0, int ncx, int ncy,
                                                        int64_t pitch, R vol,
                                                        const R* __restrict__ reg, R* h, R* hu,
                                                        R* hv, const int* err, int orient,
                                                        int cnx, int cny)
{
    if (err && *err) return;
    reflux_face(ci0, cj0, ncx, ncy, pitch, vol, reg, h, hu, hv, orient, cnx, cny,
                (int)(blockIdx.x * kBlock + threadIdx.x));
}

// Covered coarse cells from the fine cells: c = sum of the 2 x 2 children
// (jref outer, iref inner) times 1/4, AMReX average_down for cell data.
template <typename R>
__device__ __forceinline__ void avgdown_cell(int ci0, int cj0, int ncx, int ncy, int64_t fpitch,
                                             const R* fh, const R* fhu, const R* fhv,
                                             int64_t cpitch, R* h, R* hu, R* hv, int q)
{
    if (q >= ncx * ncy) return;
    const int ic = q % ncx, jc = q / ncx;
    const R* src[3] = {fh, fhu, fhv};
    R* dst[3] = {h, hu, hv};
    for (int k = 0; k < 3; ++k) {
        R c = R(0);
        for (int jr = 0; jr < 2; ++jr)
            for (int ir = 0; ir < 2; ++ir) c += src[k][(int64_t)(2 * jc + jr) * fpitch + 2 * ic + ir];
        dst[k][(int64_t)(cj0 + jc) * cpitch + ci0 + ic] = R(0.25) * c;
    }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void avgdown_state_kernel(int ci0, int cj0, int ncx, int ncy,
                                                               int64_t fpitch, const R* fh,
                                                               const R* fhu, const R* fhv,
                                                               int64_t cpitch, R* h, R* hu, R* hv,
                                                               const int* err)
{
    if (err && *err) return;
    avgdown_cell(ci0, cj0, ncx, ncy, fpitch, fh, fhu, fhv, cpitch, h, hu, hv,
                 (int)(blockIdx.x * kBlock + threadIdx.x));
}

// ---------------------------------------------------------------------------
// AMR patch-table launches (AmrPatch): one launch per operation and level
// over every box of the level.  A workgroup finds its box by binary search
// over the box's first block and does what that box's own launch does, so
// the results are the per-box launches' bit for bit.
// ---------------------------------------------------------------------------

// CFL reduction of every box of a level (its valid cells) into one slot.
template <typename R>
__global__ __launch_bounds__(kBlock) void amr_cfl_kernel(AmrLevelArgs<R> L,
                                                         const AmrPatch* __restrict__ tab,
                                                         int npatch, int mul, CflSlot* slot)
{
    const int b = blockIdx.x;
    const int k = amr_find(tab, npatch, b, [](const AmrPatch& t) { return t.cfl0; });
    const AmrPatch P = tab[k];
    const int lb = b - P.cfl0;
    cfl_block(amr_view(L, L.u_in, P.i0, P.j0, P.w, P.h), L.dx, L.dy, L.rdx, L.rdy, mul, slot,
              lb % P.cflx, lb / P.cflx);
}

// Ghost rings of every box of fine level F from coarse level C (states co
// and cn blended by alpha); cells the fine level owns keep their values.
template <typename R>
__global__ __launch_bounds__(kBlock) void amr_fill_kernel(AmrLevelArgs<R> C, const R* co,
                                                          const R* cn, R alpha, AmrLevelArgs<R> F,
                                                          const AmrPatch* __restrict__ tab,
                                                          int npatch)
{
    const int b = blockIdx.x;
    const int k = amr_find(tab, npatch, b, [](const AmrPatch& t) { return t.fill0; });
    const AmrPatch P = tab[k];
    const PatchView<R> po = amr_view(C, co, 0, 0, C.nx, C.ny);
    const PatchView<R> pn = amr_view(C, cn, 0, 0, C.nx, C.ny);
    const int64_t off = (int64_t)(P.j0 + 2) * F.pitch + (P.i0 + 2);
    R* u = F.u_out;
    fillfine_cell(po, pn, alpha, P.i0 / 2, P.j0 / 2, P.w, P.h, F.pitch, 0, F.terrain + off, u + off,
                  u + F.plane + off, u + 2 * F.plane + off, F.mask + off, F.pitch,
                  (b - P.fill0) * kBlock + (int)threadIdx.x);
}

// Reflux of one side (orient) of every box of fine table tab into coarse
// level C's current state (C.u_out); regs as amr_faces_kernel's.
template <typename R>
__global__ __launch_bounds__(kBlock) void amr_reflux_kernel(AmrLevelArgs<R> C,
                                                            const AmrPatch* __restrict__ tab,
                                                            int npatch, const R* __restrict__ regs,
                                                            int orient, R vol)
{
    const int b = blockIdx.x;
    const int k = amr_find(tab, npatch, b, [](const AmrPatch& t) { return t.face0; });
    const AmrPatch P = tab[k];
    const int64_t off = 2 * C.pitch + 2;
    R* u = C.u_out;
    reflux_face(P.i0 / 2, P.j0 / 2, P.w / 2, P.h / 2, C.pitch, vol, regs + P.reg, u + off,
                u + C.plane + off, u + 2 * C.plane + off, orient, C.nx, C.ny,
                (b - P.face0) * kBlock + (int)threadIdx.x);
}

// average_down of every box of fine level F (state F.u_in) into coarse
// level C's current state (C.u_out).
template <typename R>
__global__ __launch_bounds__(kBlock) void amr_avgdown_kernel(AmrLevelArgs<R> F, AmrLevelArgs<R> C,
                                                             const AmrPatch* __restrict__ tab,
                                                             int npatch)
{
    const int b = blockIdx.x;
    const int k = amr_find(tab, npatch, b, [](const AmrPatch& t) { return t.avg0; });
    const AmrPatch P = tab[k];
    const int64_t fo = (int64_t)(P.j0 + 2) * F.pitch + (P.i0 + 2);
    const int64_t co = 2 * C.pitch + 2;
    const R* f = F.u_in;
    R* c = C.u_out;
    avgdown_cell(P.i0 / 2, P.j0 / 2, P.w / 2, P.h / 2, F.pitch, f + fo, f + F.plane + fo,
                 f + 2 * F.plane + fo, C.pitch, c + co, c + C.plane + co, c + 2 * C.plane + co,
                 (b - P.avg0) * kBlock + (int)threadIdx.x);
}

}  // namespace hx

// ===========================================================================
// C-ABI
// ===========================================================================
using namespace hx;
using namespace hx::host;

// Device memory the fused single-level loop works in besides the state: the
// 3 rotating CFL slots, the error flag and a dt / max|vel| history of cap
// steps (dt at hist[s], max|vel| at hist[cap + s]).
struct RunScratch {
    CflSlot* slots;
    int* err;
    double* hist;
    int cap;
};

// The allocation-free loop's scratch (hx_run_level_noalloc): one copy per
// device, held until the process ends.  Layout: 3 CflSlots at byte 0, the
// error flag at byte kScratchErr, the history of kNoAllocSteps steps at byte
// kScratchHist.  Host code never uses this name as an address: hx_ctx_create
// resolves the device copy's address with hipGetSymbolAddress once per device.
constexpr int kNoAllocSteps = 8192;
constexpr size_t kScratchErr = 3 * sizeof(CflSlot);
constexpr size_t kScratchHist = kScratchErr + 256;
constexpr size_t kScratchBytes = kScratchHist + sizeof(double) * 2 * kNoAllocSteps;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];       // device address of scratch, per device
static std::mutex g_scratch_mu;

static RunScratch scratch_run(void* base)
{
    char* b = (char*)base;
    return RunScratch{(CflSlot*)b, (int*)(b + kScratchErr), (double*)(b + kScratchHist), kNoAllocSteps};
}

// Launch-shape defaults from the strip-height / occupancy sweep on MI355X
// (4096^2 plain Roe and 2048^2 hydrostatic workloads): short strips keep
// enough wavefronts in flight; the FP64 kernels run best when the register
// allocator leaves room for 3 waves per SIMD.
extern "C" int hx_abi_version(void) { return HX_ABI_VERSION; }

extern "C" const char* hx_status_string(int s)
{
    switch (s) {
        case HX_OK: return "ok";
        case HX_E_BADARG: return "invalid argument";
        case HX_E_DT_COLLAPSE: return "dt collapsed below 1e-12";
        case HX_E_HIP: return "HIP runtime error";
        case HX_E_NOMEM: return "device allocation failed";
        case HX_E_STATE: return "context used out of order";
        default: return "unknown status";
    }
}

extern "C" int hx_ctx_create(hx_ctx** out, int device)
{
    if (!out) return HX_E_BADARG;
    *out = nullptr;
    int ndev = 0;
    HX_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return HX_E_BADARG;
    HX_TRY(hipSetDevice(device));
    hx_ctx* c = new (std::nothrow) hx_ctx();
    if (!c) return HX_E_NOMEM;
    c->device = device;
    c->hist_cap = kHistCap;
    // pinned staging for the run loop's read-back: the error flag, then up to
    // the larger chunk's dt and vel histories (pageable copies took 20-30 us
    // each after a run's last step)
    const size_t pin = sizeof(double) * (1 + 2 * (size_t)std::max(kHistCap, kNoAllocSteps));
    static_assert(sizeof(double) * (1 + 2 * (size_t)kHistCap) >= sizeof(CflSlot) * HX_AMR_MAX_LEVELS,
                  "the AMR runner reads its level slots back through the pinned staging");
    if (hipMalloc(&c->slots, 4 * sizeof(CflSlot)) != hipSuccess ||
        hipMalloc(&c->err, sizeof(int)) != hipSuccess ||
        hipMalloc(&c->hist, sizeof(double) * 2 * kHistCap) != hipSuccess ||
        hipMalloc(&c->tlw, sizeof(TwoLevelWork)) != hipSuccess ||
        hipMalloc(&c->amr_slots, sizeof(CflSlot) * HX_AMR_MAX_LEVELS) != hipSuccess ||
        hipHostMalloc(&c->hpin, pin, hipHostMallocDefault) != hipSuccess) {
        hipFree(c->slots);
        hipFree(c->err);
        hipFree(c->hist);
        hipFree(c->tlw);
        hipFree(c->amr_slots);
        if (c->hpin) hipHostFree(c->hpin);
        delete c;
        return HX_E_NOMEM;
    }
    c->mark.valid = 0;
    c->amr_n = -1;
    if (device >= kMaxDevices) {
        hx_ctx_destroy(c);
        return HX_E_BADARG;
    }
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch[device]) {
            void* ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
                hx_ctx_destroy(c);
                return HX_E_HIP;
            }
            g_scratch[device] = ptr;
        }
        c->scratch = g_scratch[device];
    }
    for (int s = 0; s < 2; ++s)
        for (int d = 0; d < 2; ++d) {
            c->rows[s][d] = 0;
            c->occ[s][d] = 0;
        }
    c->arith = HX_ARITH_EXACT;
    if (hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        c->cus = 256;
    hipLaunchKernelGGL(clear_slots_kernel, dim3(1), dim3(64), 0, 0, c->slots, 4, c->err);
    HX_TRY(hipDeviceSynchronize());
    *out = c;
    return HX_OK;
}

extern "C" int hx_ctx_destroy(hx_ctx* c)
{
    if (!c) return HX_OK;
    hipSetDevice(c->device);
    hipFree(c->slots);
    hipFree(c->err);
    hipFree(c->hist);
    hipFree(c->tlw);
    hipFree(c->amr_slots);
    if (c->hpin) hipHostFree(c->hpin);
    if (c->amr_tab) hipFree(c->amr_tab);
    if (c->amr_host) hipHostFree(c->amr_host);
    if (c->pflags) hipFree(c->pflags);
    if (c->pboxes) hipFree(c->pboxes);
    if (c->cons) hipFree(c->cons);
    delete c;
    return HX_OK;
}

extern "C" int hx_set_tuning(hx_ctx* c, int solver, int dtype, int rows, int occupancy)
{
    if (!c || solver < 0 || solver > 1 || dtype < 0 || dtype > 1 || rows < 0 || rows > 1 << 20)
        return HX_E_BADARG;
    bool known = (occupancy == 0);
    for (int v = 0; v < kOccVariants; ++v) known = known || (occupancy == kOcc[v]);
    if (!known) return HX_E_BADARG;
    c->rows[solver][dtype] = rows;
    c->occ[solver][dtype] = occupancy;
    return HX_OK;
}

extern "C" int hx_set_arith(hx_ctx* c, int mode)
{
    if (!c) return HX_E_STATE;
    if (mode != HX_ARITH_EXACT && mode != HX_ARITH_FAST) return HX_E_BADARG;
    c->arith = mode;
    return HX_OK;
}

int hx::host::check_patch(const hx_patch* in, int need_ghost)
{
    if (!in || in->nx < 2 || in->ny < 2 || in->ng < 0 || in->ng > 8) return HX_E_BADARG;
    if (in->dtype != HX_F64 && in->dtype != HX_F32) return HX_E_BADARG;
    if (!in->h || !in->hu || !in->hv || !in->z || !in->n) return HX_E_BADARG;
    if (!(in->dx > 0.0) || !(in->dy > 0.0)) return HX_E_BADARG;
    for (int d = 0; d < 2; ++d) {
        const int lo = in->bc_lo[d], hi = in->bc_hi[d];
        if (lo < -1 || lo > 1 || hi < -1 || hi > 1) return HX_E_BADARG;
        if ((lo == HX_BC_GHOST || hi == HX_BC_GHOST) && in->ng < need_ghost) return HX_E_BADARG;
    }
    const int64_t gx = (in->bc_lo[0] == HX_BC_GHOST ? in->ng : 0) + (in->bc_hi[0] == HX_BC_GHOST ? in->ng : 0);
    if (in->pitch < (int64_t)in->nx + gx) return HX_E_BADARG;
    // the advance addresses a strip with 32-bit byte offsets: at this pitch
    // bound an FP64 strip of up to 3 rows fits (max_strip_rows)
    if (in->pitch > ((int64_t)1 << 26)) return HX_E_BADARG;
    return HX_OK;
}

int hx::host::advance_simd_waves(const hx_ctx* c, int solver, int dtype)
{
    const int occ = eff_occ(c, solver, dtype);
    const bool fa = (c->arith == HX_ARITH_FAST);
    if (dtype == HX_F64) return fa ? fast::simd_waves<double>(solver, occ) : exact::simd_waves<double>(solver, occ);
    return fa ? fast::simd_waves<float>(solver, occ) : exact::simd_waves<float>(solver, occ);
}

template <typename R>
static int launch(const hx_ctx* c, int solver, int dtype, const StepArgs<R>& a0, hipStream_t st)
{
    const int occ = eff_occ(c, solver, dtype);
    StepArgs<R> a = a0;
    plan_bands(a, solver, c->cus, advance_simd_waves(c, solver, dtype));
    const int rc = (c->arith == HX_ARITH_FAST) ? fast::launch_advance<R>(solver, occ, a, st)
                                               : exact::launch_advance<R>(solver, occ, a, st);
    return rc ? HX_E_HIP : HX_OK;
}

template <typename R>
static int advance_t(hx_ctx* c, int solver, const hx_patch* in, double dt, void* ho, void* huo,
                     void* hvo, int64_t opitch, void* ef, hipStream_t st)
{
    StepArgs<R> a{};
    a.p = view_of<R>(in);
    a.ho = (R*)ho;
    a.huo = (R*)huo;
    a.hvo = (R*)hvo;
    a.opitch = opitch;
    a.dt_host = dt;
    a.ef = (R*)ef;
    fill_geometry(a, solver, eff_rows(c, solver, in->dtype, in->nx, in->ny), in->dx, in->dy);
    return launch(c, solver, in->dtype, a, st);
}

static int advance_any(hx_ctx* c, int solver, const hx_patch* in, double dt, void* ho, void* huo,
                       void* hvo, int64_t opitch, void* ef, void* stream)
{
    if (!c) return HX_E_STATE;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    int rc = check_patch(in, solver == HX_SOLVER_ROE ? 1 : 2);
    if (rc) return rc;
    if (!ho || !huo || !hvo || opitch < in->nx || opitch > ((int64_t)1 << 26) || !(dt >= 0.0))
        return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    return (in->dtype == HX_F64) ? advance_t<double>(c, solver, in, dt, ho, huo, hvo, opitch, ef, st)
                                 : advance_t<float>(c, solver, in, dt, ho, huo, hvo, opitch, ef, st);
}

extern "C" int hx_advance_level(hx_ctx* c, int solver, const hx_patch* in, double dt, void* ho,
                                void* huo, void* hvo, int64_t opitch, void* stream)
{
    return advance_any(c, solver, in, dt, ho, huo, hvo, opitch, nullptr, stream);
}

extern "C" int hx_advance_level_fluxes(hx_ctx* c, int solver, const hx_patch* in, double dt,
                                       void* ho, void* huo, void* hvo, int64_t opitch,
                                       void* edge_flux, void* stream)
{
    if (!edge_flux) return HX_E_BADARG;
    return advance_any(c, solver, in, dt, ho, huo, hvo, opitch, edge_flux, stream);
}

template <typename R>
static int launch_cfl(const hx_patch* in, CflSlot* slot, hipStream_t st, int spread = 1)
{
    const int mul = is_pow2(in->dx) && is_pow2(in->dy) ? 1 : 0;
    // runs of 16 rows, shortened (to 4 at the least, a multiple of the
    // unroll) until the grid has about 1024 workgroups: a 1026^2 region took
    // 15.2 us in 325 workgroups of 16 rows
    const int64_t bx = (in->nx + kBlock - 1) / kBlock;
    int rows = kCflRows;
    while (rows > kCflUnroll && bx * ((in->ny + rows - 1) / rows) < 1024) rows -= kCflUnroll;
    const dim3 grid((unsigned)bx, (unsigned)((in->ny + rows - 1) / rows));
    hipLaunchKernelGGL(cfl_kernel<R>, grid, dim3(kBlock), 0, st, view_of<R>(in),
                       (R)in->dx, (R)in->dy, (R)(1.0 / in->dx), (R)(1.0 / in->dy), mul, slot,
                       spread, rows);
    return hip_status(hipGetLastError());
}

// Entry points for the other translation units (hx_host.hpp).
int hx::host::launch_step_f64(const hx_ctx* c, int solver, const StepArgs<double>& a,
                              hipStream_t st)
{
    return launch<double>(c, solver, HX_F64, a, st);
}

int hx::host::launch_adj_batch_f64(const hx_ctx* c, int solver, const AdjBatch& b, hipStream_t st)
{
    const int occ = eff_occ(c, solver, HX_F64);
    const int rc = (c->arith == HX_ARITH_FAST) ? fast::launch_adj_batch(solver, occ, b, st)
                                               : exact::launch_adj_batch(solver, occ, b, st);
    return rc ? HX_E_HIP : HX_OK;
}

int hx::host::reduce_cfl_f64(const hx_patch* in, CflSlot* slot, hipStream_t st)
{
    return launch_cfl<double>(in, slot, st);
}

int hx::host::clear_slots(CflSlot* s, int n, int* err, hipStream_t st)
{
    hipLaunchKernelGGL(clear_slots_kernel, dim3(1), dim3(64), 0, st, s, n, err);
    return hip_status(hipGetLastError());
}

// Grow the context's dt / max|vel| history to hold nsteps entries.
extern "C" int hx_level_dt(hx_ctx* c, int solver, const hx_patch* in, double cfl, double* dt_out,
                           double* max_vel_out, void* stream)
{
    if (!c) return HX_E_STATE;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    int rc = check_patch(in, 0);
    if (rc) return rc;
    if (!dt_out || !(cfl > 0.0)) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(clear_slots_kernel, dim3(1), dim3(64), 0, st, c->slots + 3, 1, (int*)nullptr);
    rc = (in->dtype == HX_F64) ? launch_cfl<double>(in, c->slots + 3, st)
                               : launch_cfl<float>(in, c->slots + 3, st);
    if (rc) return rc;
    CflSlot sl;
    HX_TRY(hipMemcpyAsync(c->hpin, c->slots + 3, sizeof(sl), hipMemcpyDeviceToHost, st));
    HX_TRY(hipStreamSynchronize(st));
    std::memcpy(&sl, c->hpin, sizeof(sl));
    const CflValue s = slot_value(&sl);
    double vmax;
    std::memcpy(&vmax, &s.vmax, sizeof(vmax));
    double est;
    if (in->dtype == HX_F64) {
        est = slot_dt_est<double>(s.amax, s.flags);
        *dt_out = est * cfl;
    } else {
        const float e = slot_dt_est<float>(s.amax, s.flags);
        est = e;
        *dt_out = (double)(e * (float)cfl);
    }
    if (max_vel_out) *max_vel_out = (s.flags ? vmax : -1.7976931348623157e308);
    return (est < 1.e-12) ? HX_E_DT_COLLAPSE : HX_OK;
}

extern "C" int hx_tag_level(hx_ctx* c, int solver, const hx_patch* in, double time,
                            const double* thresh, uint8_t* tags, void* stream)
{
    if (!c) return HX_E_STATE;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    int rc = check_patch(in, 1);
    if (rc) return rc;
    if (!thresh || !tags) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((in->nx + 63) / 64, (in->ny + 3) / 4);
    if (in->dtype == HX_F64)
        hipLaunchKernelGGL(tag_kernel<double>, grid, dim3(256), 0, st, view_of<double>(in), solver,
                           in->dx, in->dy, time, thresh[0], thresh[1], thresh[2], tags);
    else
        hipLaunchKernelGGL(tag_kernel<float>, grid, dim3(256), 0, st, view_of<float>(in), solver,
                           (float)in->dx, (float)in->dy, time, (float)thresh[0], (float)thresh[1],
                           (float)thresh[2], tags);
    return hip_status(hipGetLastError());
}

extern "C" int hx_masked_average_down(hx_ctx* c, int dtype, int nxf, int nyf, int ncomp,
                                      const void* fine, int rx, int ry, double dxf, double dyf,
                                      double nodata, void* crse, int32_t* nvalid, void* stream)
{
    if (!c) return HX_E_STATE;
    if ((dtype != HX_F64 && dtype != HX_F32) || !fine || !crse || nxf < 1 || nyf < 1 ||
        ncomp < 1 || rx < 1 || ry < 1 || nxf % rx || nyf % ry)
        return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = (int64_t)(nxf / rx) * (nyf / ry) * ncomp;
    const unsigned nblk = (unsigned)((n + 255) / 256);
    if (dtype == HX_F64)
        hipLaunchKernelGGL(avgdown_kernel<double>, dim3(nblk), dim3(256), 0, st, (const double*)fine,
                           nxf, nyf, ncomp, rx, ry, dxf * dyf, nodata, (double*)crse, nvalid);
    else
        hipLaunchKernelGGL(avgdown_kernel<float>, dim3(nblk), dim3(256), 0, st, (const float*)fine,
                           nxf, nyf, ncomp, rx, ry, (float)dxf * (float)dyf, (float)nodata,
                           (float*)crse, nvalid);
    return hip_status(hipGetLastError());
}

extern "C" int hx_prune_boxes(hx_ctx* c, int dtype, int nbox, const int32_t* boxes, int nxs,
                              int nys, const void* z, int scale_log2, uint8_t* keep,
                              uint8_t* has_nodata, int32_t* nkeep, void* stream)
{
    if (!c) return HX_E_STATE;
    if ((dtype != HX_F64 && dtype != HX_F32) || nbox < 0 || (nbox > 0 && !boxes) || !z ||
        nxs < 1 || nys < 1 || scale_log2 < 0 || scale_log2 > 20 || !keep || !has_nodata || !nkeep)
        return HX_E_BADARG;
    *nkeep = 0;
    if (nbox == 0) return HX_OK;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    if (nbox > c->pcap) {
        if (c->pflags) hipFree(c->pflags);
        if (c->pboxes) hipFree(c->pboxes);
        c->pflags = nullptr;
        c->pboxes = nullptr;
        c->pcap = 0;
        if (hipMalloc(&c->pflags, sizeof(uint32_t) * 2 * nbox) != hipSuccess ||
            hipMalloc(&c->pboxes, sizeof(int32_t) * 4 * nbox) != hipSuccess)
            return HX_E_NOMEM;
        c->pcap = nbox;
    }
    HX_TRY(hipMemcpyAsync(c->pboxes, boxes, sizeof(int32_t) * 4 * nbox, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(zero_u32_kernel, dim3((2 * nbox + 255) / 256), dim3(256), 0, st, c->pflags, 2 * nbox);
    if (dtype == HX_F64)
        hipLaunchKernelGGL(prune_kernel<double>, dim3(nbox), dim3(256), 0, st, c->pboxes,
                           (const double*)z, nxs, nys, 1 << scale_log2, c->pflags);
    else
        hipLaunchKernelGGL(prune_kernel<float>, dim3(nbox), dim3(256), 0, st, c->pboxes,
                           (const float*)z, nxs, nys, 1 << scale_log2, c->pflags);
    HX_TRY(hipGetLastError());
    uint32_t* hf = (uint32_t*)std::malloc(sizeof(uint32_t) * 2 * nbox);
    if (!hf) return HX_E_NOMEM;
    hipError_t e = hipMemcpyAsync(hf, c->pflags, sizeof(uint32_t) * 2 * nbox, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { std::free(hf); return HX_E_HIP; }
    int nk = 0;
    for (int b = 0; b < nbox; ++b) {
        keep[b] = (uint8_t)(hf[2 * b] ? 1 : 0);
        has_nodata[b] = (uint8_t)(hf[2 * b + 1] ? 1 : 0);
        nk += keep[b];
    }
    std::free(hf);
    if (nk == 0) {   // AmrMeshState_AMR.cpp:397-401: keep the whole layout
        for (int b = 0; b < nbox; ++b) { keep[b] = 1; has_nodata[b] = 0; }
        nk = nbox;
    }
    *nkeep = nk;
    return HX_OK;
}

template <typename R>
static int cons_t(hx_ctx* c, const hx_patch* in, hipStream_t st)
{
    const int64_t ncell = (int64_t)in->nx * in->ny;
    const int nblk = (int)std::min<int64_t>(kConsBlocks, (ncell + kBlock - 1) / kBlock);
    const R dA = (R)in->dx * (R)in->dy;
    hipLaunchKernelGGL(cons_kernel<R>, dim3(nblk), dim3(kBlock), 0, st, view_of<R>(in), dA,
                       c->cons);
    hipLaunchKernelGGL(cons_final_kernel, dim3(1), dim3(kConsBlocks), 0, st, (const double*)c->cons, nblk,
                       c->cons + (size_t)kConsBlocks * 9);
    return hip_status(hipGetLastError());
}

extern "C" int hx_conservation(hx_ctx* c, const hx_patch* in, double* out, void* stream)
{
    if (!c) return HX_E_STATE;
    int rc = check_patch(in, 0);
    if (rc) return rc;
    if (!out) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    if (!c->cons && hipMalloc(&c->cons, sizeof(double) * ((size_t)kConsBlocks * 9 + 5)) != hipSuccess) {
        c->cons = nullptr;
        return HX_E_NOMEM;
    }
    hipStream_t st = (hipStream_t)stream;
    rc = (in->dtype == HX_F64) ? cons_t<double>(c, in, st) : cons_t<float>(c, in, st);
    if (rc) return rc;
    HX_TRY(hipMemcpyAsync(out, c->cons + (size_t)kConsBlocks * 9, sizeof(double) * 5,
                          hipMemcpyDeviceToHost, st));
    HX_TRY(hipStreamSynchronize(st));
    return HX_OK;
}

// Streaming ceiling of the advance's access pattern: an elementwise pass that
// reads five FP64 planes and writes three (x = a + s, y = b s, z = c - s with
// s = d + e), as coalesced and as simple as a kernel with this traffic can
// be, with the fastest launch measured for it on MI355X (non-temporal loads,
// 512 workgroups of 256 threads; 1024 and more workgroups, or cached loads,
// were 3-8 % slower).  Its time bounds from below what an advance of the
// same planes can reach on the device; bench.py reports the advance
// against it.
__global__ __launch_bounds__(256) void stream5r3w_kernel(const double* __restrict__ a,
                                                        const double* __restrict__ b,
                                                        const double* __restrict__ c,
                                                        const double* __restrict__ d,
                                                        const double* __restrict__ e,
                                                        double* __restrict__ x, double* __restrict__ y,
                                                        double* __restrict__ z, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double sv = __builtin_nontemporal_load(d + i) + __builtin_nontemporal_load(e + i);
        x[i] = __builtin_nontemporal_load(a + i) + sv;
        y[i] = __builtin_nontemporal_load(b + i) * sv;
        z[i] = __builtin_nontemporal_load(c + i) - sv;
    }
}

extern "C" int hx_stream_probe(hx_ctx* c, int64_t n, const void* const* in, void* const* out,
                               void* stream)
{
    if (!c) return HX_E_STATE;
    if (n <= 0 || !in || !out) return HX_E_BADARG;
    for (int k = 0; k < 5; ++k)
        if (!in[k]) return HX_E_BADARG;
    for (int k = 0; k < 3; ++k)
        if (!out[k]) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 512);
    hipLaunchKernelGGL(stream5r3w_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const double*)in[0], (const double*)in[1], (const double*)in[2],
                       (const double*)in[3], (const double*)in[4], (double*)out[0], (double*)out[1],
                       (double*)out[2], n);
    return hipGetLastError() == hipSuccess ? HX_OK : HX_E_HIP;
}

// Steps [s0, s1) of a rollout whose step 0 reads slot `base`: step s reads
// slot (base + s) % 3, reduces its output into the next slot and clears the
// one after.  History and error indices are relative to s0 (the context's
// history holds kHistCap steps).  When s0 == 0 and fresh, the slots and the
// error flag are reset and the initial state's CFL is reduced into `base`.
template <typename R>
static int run_t(hx_ctx* c, int solver, const hx_patch* in, void* ah, void* ahu, void* ahv,
                 double cfl, int s0, int s1, int base, bool fresh, const RunScratch& sc,
                 hipStream_t st)
{
    StepArgs<R> a{};
    a.p = view_of<R>(in);
    a.opitch = in->pitch;
    a.cfl = cfl;
    a.err = sc.err;
    a.dt_hist = sc.hist;
    a.vel_hist = sc.hist + sc.cap;
    fill_geometry(a, solver, eff_rows(c, solver, in->dtype, in->nx, in->ny), in->dx, in->dy);
    // every step reads the slot the step before reduced into: both sides use
    // the level's rule (the initial reduction too, and the steps of a resumed
    // run, which take the same level)
    a.cfl_spread = cfl_spread_for(in->nx, in->ny);
    if (s0 == 0 && fresh) {
        hipLaunchKernelGGL(clear_slots_kernel, dim3(1), dim3(64), 0, st, sc.slots, 3, sc.err);
        const int rc = launch_cfl<R>(in, sc.slots + base, st, a.cfl_spread);
        if (rc) return rc;
    }
    R* bufA[3] = {(R*)in->h, (R*)in->hu, (R*)in->hv};
    R* bufB[3] = {(R*)ah, (R*)ahu, (R*)ahv};
    for (int s = s0; s < s1; ++s) {
        R** src = (s & 1) ? bufB : bufA;
        R** dst = (s & 1) ? bufA : bufB;
        a.p.h = src[0];
        a.p.hu = src[1];
        a.p.hv = src[2];
        a.ho = dst[0];
        a.huo = dst[1];
        a.hvo = dst[2];
        a.rd = sc.slots + ((base + s) % 3);
        a.acc = sc.slots + ((base + s + 1) % 3);
        a.clr = sc.slots + ((base + s + 2) % 3);
        a.step = s - s0;
        const int rc = launch(c, solver, in->dtype, a, st);
        if (rc) return rc;
    }
    return HX_OK;
}

static void set_mark(hx_ctx* c, int solver, const hx_patch* in, void* ah, void* ahu, void* ahv,
                     int nsteps, int base)
{
    RunMark& m = c->mark;
    m.valid = 1;
    m.solver = solver;
    m.dtype = in->dtype;
    m.nx = in->nx;
    m.ny = in->ny;
    m.pitch = in->pitch;
    m.dx = in->dx;
    m.dy = in->dy;
    m.bc[0] = in->bc_lo[0]; m.bc[1] = in->bc_lo[1];
    m.bc[2] = in->bc_hi[0]; m.bc[3] = in->bc_hi[1];
    const bool odd = nsteps & 1;
    m.h = odd ? ah : in->h;
    m.hu = odd ? ahu : in->hu;
    m.hv = odd ? ahv : in->hv;
    m.z = in->z;
    m.n = in->n;
    m.slot = (base + nsteps) % 3;
}

static bool mark_matches(const hx_ctx* c, int solver, const hx_patch* in)
{
    const RunMark& m = c->mark;
    return m.valid && m.solver == solver && m.dtype == in->dtype && m.nx == in->nx &&
           m.ny == in->ny && m.pitch == in->pitch && m.dx == in->dx && m.dy == in->dy &&
           m.bc[0] == in->bc_lo[0] && m.bc[1] == in->bc_lo[1] && m.bc[2] == in->bc_hi[0] &&
           m.bc[3] == in->bc_hi[1] && m.h == in->h && m.hu == in->hu && m.hv == in->hv &&
           m.z == in->z && m.n == in->n;
}

// noalloc: the loop works in the device's static scratch instead of the
// context's buffers (at most kNoAllocSteps steps, no resume mark).
static int run_level(hx_ctx* c, int solver, const hx_patch* in, void* ah, void* ahu, void* ahv,
                     double cfl, int nsteps, double* dt_hist, double* vel_hist,
                     int32_t* steps_done, void* stream, bool resume, bool noalloc = false)
{
    if (!c) return HX_E_STATE;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    int rc = check_patch(in, 0);
    if (rc) return rc;
    for (int d = 0; d < 2; ++d)
        if (in->bc_lo[d] == HX_BC_GHOST || in->bc_hi[d] == HX_BC_GHOST) return HX_E_BADARG;
    if (in->eta || !ah || !ahu || !ahv || nsteps < 0 || !(cfl > 0.0)) return HX_E_BADARG;
    if (ah == in->h || ahu == in->hu || ahv == in->hv) return HX_E_BADARG;
    if (resume && !mark_matches(c, solver, in)) return HX_E_STATE;
    if (noalloc && !c->scratch) return HX_E_STATE;
    if (noalloc && nsteps > kNoAllocSteps) return HX_E_BADARG;
    if (steps_done) *steps_done = 0;
    if (nsteps == 0) return HX_OK;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    const RunScratch sc = noalloc ? scratch_run(c->scratch)
                                  : RunScratch{c->slots, c->err, c->hist, c->hist_cap};
    const int base = resume ? c->mark.slot : 0;
    if (!noalloc) c->mark.valid = 0;
    for (int s0 = 0; s0 < nsteps; s0 += sc.cap) {
        const int s1 = std::min(nsteps, s0 + sc.cap);
        rc = (in->dtype == HX_F64)
                 ? run_t<double>(c, solver, in, ah, ahu, ahv, cfl, s0, s1, base, !resume, sc, st)
                 : run_t<float>(c, solver, in, ah, ahu, ahv, cfl, s0, s1, base, !resume, sc, st);
        if (rc) return rc;
        // read back through the context's pinned staging: [error flag][dt][vel]
        const size_t n = (size_t)(s1 - s0);
        int* perr = (int*)c->hpin;
        double* pdt = c->hpin + 1;
        double* pvel = pdt + n;
        HX_TRY(hipMemcpyAsync(perr, sc.err, sizeof(int), hipMemcpyDeviceToHost, st));
        if (dt_hist)
            HX_TRY(hipMemcpyAsync(pdt, sc.hist, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        if (vel_hist)
            HX_TRY(hipMemcpyAsync(pvel, sc.hist + sc.cap, sizeof(double) * n,
                                  hipMemcpyDeviceToHost, st));
        HX_TRY(hipStreamSynchronize(st));
        const int err = *perr;
        if (dt_hist) std::memcpy(dt_hist + s0, pdt, sizeof(double) * n);
        if (vel_hist) std::memcpy(vel_hist + s0, pvel, sizeof(double) * n);
        if (err) {
            if (steps_done) *steps_done = s0 + err - 1;
            return HX_E_DT_COLLAPSE;
        }
    }
    if (!noalloc) set_mark(c, solver, in, ah, ahu, ahv, nsteps, base);
    if (steps_done) *steps_done = nsteps;
    return HX_OK;
}

extern "C" int hx_run_level(hx_ctx* c, int solver, const hx_patch* in, void* ah, void* ahu,
                            void* ahv, double cfl, int nsteps, double* dt_hist, double* vel_hist,
                            int32_t* steps_done, void* stream)
{
    return run_level(c, solver, in, ah, ahu, ahv, cfl, nsteps, dt_hist, vel_hist, steps_done,
                     stream, false);
}

extern "C" int hx_run_level_noalloc(hx_ctx* c, int solver, const hx_patch* in, void* ah,
                                    void* ahu, void* ahv, double cfl, int nsteps,
                                    double* dt_hist, double* vel_hist, int32_t* steps_done,
                                    void* stream)
{
    return run_level(c, solver, in, ah, ahu, ahv, cfl, nsteps, dt_hist, vel_hist, steps_done,
                     stream, false, true);
}

extern "C" int hx_run_level_resume(hx_ctx* c, int solver, const hx_patch* in, void* ah, void* ahu,
                                   void* ahv, double cfl, int nsteps, double* dt_hist,
                                   double* vel_hist, int32_t* steps_done, void* stream)
{
    return run_level(c, solver, in, ah, ahu, ahv, cfl, nsteps, dt_hist, vel_hist, steps_done,
                     stream, true);
}

// ---------------------------------------------------------------------------
// Batched ensemble
// ---------------------------------------------------------------------------
struct hx_ensemble {
    hx_ctx* ctx;
    int solver, dtype, nscen;
    int rows, occ, persistent;
    int total;
    std::vector<hx_patch> patches;
    std::vector<void*> alt;      // 3 per scenario: buffer B
    void* desc;                  // device ScenDesc<R>[nscen]
    int* prefix;                 // device [nscen + 1]
    CflSlot* slots;              // device [3 * nscen]
    int* err;                    // device [nscen]
    double* hist;                // device [2 * nscen * cap]
    int cap;
    unsigned* queue;             // device [3]
    int* step_base;              // device [1]
    int ncu;
};

template <typename R>
static int ens_upload(hx_ensemble* e)
{
    const int cols = (e->solver == HX_SOLVER_ROE) ? kRoeCols : kHydroCols;
    std::vector<ScenDesc<R>> d(e->nscen);
    std::vector<int> pre(e->nscen + 1, 0);
    // table order: largest scenarios first, so the long strips of big levels
    // start early and small levels fill the tail of the launch
    std::vector<int> order(e->nscen);
    for (int s = 0; s < e->nscen; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (int64_t)e->patches[a].nx * e->patches[a].ny > (int64_t)e->patches[b].nx * e->patches[b].ny;
    });
    int total = 0;
    for (int t = 0; t < e->nscen; ++t) {
        const int s = order[t];
        const hx_patch* in = &e->patches[s];
        // strips are addressed with 32-bit byte offsets (max_strip_rows)
        if (e->rows > max_strip_rows<R>(in->pitch, in->pitch)) return HX_E_BADARG;
        ScenDesc<R>& x = d[t];
        std::memset(&x, 0, sizeof(x));
        x.p = view_of<R>(in);
        x.bh = (R*)e->alt[3 * s];
        x.bhu = (R*)e->alt[3 * s + 1];
        x.bhv = (R*)e->alt[3 * s + 2];
        x.dx = (R)in->dx;
        x.dy = (R)in->dy;
        x.rdx = (R)(1.0 / in->dx);
        x.rdy = (R)(1.0 / in->dy);
        x.dx_pow2 = is_pow2(in->dx) && is_pow2(in->dy) ? 1 : 0;
        x.nsx = (in->nx + cols - 1) / cols;
        x.nstrips = x.nsx * ((in->ny + e->rows - 1) / e->rows);
        x.strip0 = total;
        x.slots = e->slots + 3 * s;
        x.err = e->err + s;
        x.dt_hist = e->hist + (size_t)s * e->cap;
        x.vel_hist = e->hist + (size_t)e->nscen * e->cap + (size_t)s * e->cap;
        pre[t] = total;
        total += x.nstrips;
    }
    pre[e->nscen] = total;
    e->total = total;
    HX_TRY(hipMemcpy(e->desc, d.data(), sizeof(ScenDesc<R>) * e->nscen, hipMemcpyHostToDevice));
    HX_TRY(hipMemcpy(e->prefix, pre.data(), sizeof(int) * (e->nscen + 1), hipMemcpyHostToDevice));
    return HX_OK;
}

static int ens_upload_any(hx_ensemble* e)
{
    return (e->dtype == HX_F64) ? ens_upload<double>(e) : ens_upload<float>(e);
}

static void ens_free(hx_ensemble* e)
{
    if (!e) return;
    hipFree(e->desc);
    hipFree(e->prefix);
    hipFree(e->slots);
    hipFree(e->err);
    if (e->hist) hipFree(e->hist);
    hipFree(e->queue);
    hipFree(e->step_base);
    delete e;
}

// Default strip height of an ensemble launch: the context's tuning, else one
// height for the strips of all its scenarios together (default_rows_multi),
// no taller than any scenario's offsets allow.
static int ens_default_rows(const hx_ensemble* e)
{
    std::vector<int> nx(e->nscen), ny(e->nscen);
    int64_t cap = INT64_MAX;
    for (int s = 0; s < e->nscen; ++s) {
        const hx_patch& p = e->patches[s];
        nx[s] = p.nx;
        ny[s] = p.ny;
        cap = std::min<int64_t>(cap, e->dtype == HX_F64 ? max_strip_rows<double>(p.pitch, p.pitch)
                                                        : max_strip_rows<float>(p.pitch, p.pitch));
    }
    const int r = eff_rows_multi(e->ctx, e->solver, e->dtype, nx.data(), ny.data(), e->nscen);
    return (int)std::max<int64_t>(1, std::min<int64_t>(r, cap));
}

extern "C" int hx_ensemble_create(hx_ctx* c, int solver, int dtype, int nscen, const hx_patch* scen,
                                  void* const* alt, hx_ensemble** out)
{
    if (!c) return HX_E_STATE;
    if (!out || !scen || !alt || nscen < 1) return HX_E_BADARG;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    if (dtype != HX_F64 && dtype != HX_F32) return HX_E_BADARG;
    *out = nullptr;
    for (int s = 0; s < nscen; ++s) {
        const hx_patch* in = &scen[s];
        int rc = check_patch(in, 0);
        if (rc) return rc;
        if (in->dtype != dtype || in->eta) return HX_E_BADARG;
        for (int d = 0; d < 2; ++d)
            if (in->bc_lo[d] == HX_BC_GHOST || in->bc_hi[d] == HX_BC_GHOST) return HX_E_BADARG;
        if (!alt[3 * s] || !alt[3 * s + 1] || !alt[3 * s + 2]) return HX_E_BADARG;
    }
    HX_TRY(hipSetDevice(c->device));
    hx_ensemble* e = new (std::nothrow) hx_ensemble();
    if (!e) return HX_E_NOMEM;
    e->ctx = c;
    e->solver = solver;
    e->dtype = dtype;
    e->nscen = nscen;
    e->occ = eff_occ(c, solver, dtype);
    e->persistent = 0;
    e->patches.assign(scen, scen + nscen);
    e->rows = ens_default_rows(e);
    e->alt.assign(alt, alt + 3 * nscen);
    const size_t dsz = (dtype == HX_F64) ? sizeof(ScenDesc<double>) : sizeof(ScenDesc<float>);
    e->cap = 1024;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
        ncu = 256;
    e->ncu = ncu;
    if (hipMalloc(&e->desc, dsz * nscen) != hipSuccess ||
        hipMalloc(&e->prefix, sizeof(int) * (nscen + 1)) != hipSuccess ||
        hipMalloc(&e->slots, sizeof(CflSlot) * 3 * nscen) != hipSuccess ||
        hipMalloc(&e->err, sizeof(int) * nscen) != hipSuccess ||
        hipMalloc(&e->hist, sizeof(double) * 2 * nscen * (size_t)e->cap) != hipSuccess ||
        hipMalloc(&e->queue, sizeof(unsigned) * 3) != hipSuccess ||
        hipMalloc(&e->step_base, sizeof(int)) != hipSuccess) {
        ens_free(e);
        return HX_E_NOMEM;
    }
    int rc = ens_upload_any(e);
    if (rc) { ens_free(e); return rc; }
    *out = e;
    return HX_OK;
}

extern "C" int hx_ensemble_destroy(hx_ensemble* e)
{
    if (!e) return HX_OK;
    hipSetDevice(e->ctx->device);
    ens_free(e);
    return HX_OK;
}

extern "C" int hx_ensemble_set_tuning(hx_ensemble* e, int rows, int occupancy, int persistent)
{
    if (!e) return HX_E_STATE;
    bool known = (occupancy == 0);
    for (int v = 0; v < kOccVariants; ++v) known = known || (occupancy == kOcc[v]);
    if (!known || rows < 0 || rows > 1 << 20 || (persistent != 0 && persistent != 1))
        return HX_E_BADARG;
    HX_TRY(hipSetDevice(e->ctx->device));
    e->rows = rows ? rows : ens_default_rows(e);
    e->occ = occupancy ? occupancy : default_occ(e->solver, e->dtype);
    e->persistent = persistent;
    return ens_upload_any(e);
}

template <typename R>
static int ens_step(hx_ensemble* e, double cfl, int parity, int ring, int step_off, hipStream_t st)
{
    EnsArgs<R> a{};
    a.desc = (const ScenDesc<R>*)e->desc;
    a.strip_start = e->prefix;
    a.nscen = e->nscen;
    a.total = e->total;
    a.rows = e->rows;
    a.parity = parity;
    a.ring = ring;
    a.step_base = e->step_base;
    a.step_off = step_off;
    a.cfl = cfl;
    a.queue = e->queue;
    a.persistent = e->persistent;
    const int grid = e->persistent ? std::max(1, e->ncu * e->occ) : (e->total + kWaves - 1) / kWaves;
    const int rc = (e->ctx->arith == HX_ARITH_FAST)
                       ? fast::launch_ensemble<R>(e->solver, e->occ, a, grid, st)
                       : exact::launch_ensemble<R>(e->solver, e->occ, a, grid, st);
    return rc ? HX_E_HIP : HX_OK;
}

template <typename R>
static int ens_run_t(hx_ensemble* e, double cfl, int nsteps, int use_graph, hipStream_t st)
{
    const int nsl = 3 * e->nscen;
    const int nthr = std::max(nsl, std::max(e->nscen, 3));
    hipLaunchKernelGGL(ens_reset_kernel, dim3((nthr + 255) / 256), dim3(256), 0, st, e->slots, nsl,
                       e->err, e->nscen, e->queue, e->step_base);
    hipLaunchKernelGGL(ens_cfl_kernel<R>, dim3(64, e->nscen), dim3(kBlock), 0, st,
                       (const ScenDesc<R>*)e->desc);
    HX_TRY(hipGetLastError());
    int done = 0;
    const int cycles = (use_graph && st != nullptr) ? nsteps / 6 : 0;
    if (cycles > 0) {
        // one 6-step cycle (2 state buffers x 3 CFL slots) captured once and
        // replayed; the tick node advances the device step counter
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HX_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = HX_OK;
        for (int k = 0; k < 6 && rc == HX_OK; ++k) rc = ens_step<R>(e, cfl, k & 1, k % 3, k, st);
        if (rc == HX_OK) hipLaunchKernelGGL(ens_tick_kernel, dim3(1), dim3(64), 0, st, e->step_base, 6);
        const hipError_t ce = hipStreamEndCapture(st, &g);
        if (rc != HX_OK || ce != hipSuccess) { if (g) hipGraphDestroy(g); return HX_E_HIP; }
        if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) { hipGraphDestroy(g); return HX_E_HIP; }
        hipError_t le = hipSuccess;
        for (int c = 0; c < cycles && le == hipSuccess; ++c) le = hipGraphLaunch(ge, st);
        hipGraphExecDestroy(ge);
        hipGraphDestroy(g);
        if (le != hipSuccess) return HX_E_HIP;
        done = 6 * cycles;
    }
    for (int j = 0; done + j < nsteps; ++j) {
        const int rc = ens_step<R>(e, cfl, j & 1, j % 3, (cycles > 0) ? j : done + j, st);
        if (rc) return rc;
    }
    return HX_OK;
}

extern "C" int hx_ensemble_run(hx_ensemble* e, double cfl, int nsteps, int use_graph,
                               double* dt_hist, int32_t* steps_done, void* stream)
{
    if (!e) return HX_E_STATE;
    if (nsteps < 0 || !(cfl > 0.0)) return HX_E_BADARG;
    HX_TRY(hipSetDevice(e->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    if (nsteps > e->cap) {
        int cap = e->cap;
        while (cap < nsteps) cap *= 2;
        HX_TRY(hipStreamSynchronize(st));
        hipFree(e->hist);
        e->hist = nullptr;
        if (hipMalloc(&e->hist, sizeof(double) * 2 * e->nscen * (size_t)cap) != hipSuccess) return HX_E_NOMEM;
        e->cap = cap;
        int rc = ens_upload_any(e);
        if (rc) return rc;
    }
    if (nsteps == 0) return HX_OK;
    int rc = (e->dtype == HX_F64) ? ens_run_t<double>(e, cfl, nsteps, use_graph, st)
                                  : ens_run_t<float>(e, cfl, nsteps, use_graph, st);
    if (rc) return rc;
    std::vector<int> err(e->nscen);
    HX_TRY(hipMemcpyAsync(err.data(), e->err, sizeof(int) * e->nscen, hipMemcpyDeviceToHost, st));
    if (dt_hist)
        for (int s = 0; s < e->nscen; ++s)
            HX_TRY(hipMemcpyAsync(dt_hist + (size_t)s * nsteps, e->hist + (size_t)s * e->cap,
                                  sizeof(double) * nsteps, hipMemcpyDeviceToHost, st));
    HX_TRY(hipStreamSynchronize(st));
    bool any = false;
    for (int s = 0; s < e->nscen; ++s) {
        if (steps_done) steps_done[s] = err[s] ? err[s] - 1 : nsteps;
        any = any || err[s];
    }
    return any ? HX_E_DT_COLLAPSE : HX_OK;
}

// ---------------------------------------------------------------------------
// Two-level coupling entry points
// ---------------------------------------------------------------------------
static int launch_faces(const hx_ctx* c, int solver, const hx_patch* in, const FaceReq& q,
                        void* out, hipStream_t st)
{
    int rc;
    if (in->dtype == HX_F64) {
        const PatchView<double> v = view_of<double>(in);
        rc = (c->arith == HX_ARITH_FAST)
                 ? fast::launch_bflux<double>(solver, v, in->dx, in->dy, q, (double*)out, st)
                 : exact::launch_bflux<double>(solver, v, in->dx, in->dy, q, (double*)out, st);
    } else {
        const PatchView<float> v = view_of<float>(in);
        rc = (c->arith == HX_ARITH_FAST)
                 ? fast::launch_bflux<float>(solver, v, (float)in->dx, (float)in->dy, q,
                                            (float*)out, st)
                 : exact::launch_bflux<float>(solver, v, (float)in->dx, (float)in->dy, q,
                                             (float*)out, st);
    }
    return rc ? HX_E_HIP : HX_OK;
}

extern "C" int hx_face_fluxes(hx_ctx* c, int solver, const hx_patch* in, int nfaces,
                              const int32_t* faces, int pair, double scale_x, double scale_y,
                              int accumulate, void* out, void* stream)
{
    if (!c) return HX_E_STATE;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    int rc = check_patch(in, solver == HX_SOLVER_ROE ? 1 : 2);
    if (rc) return rc;
    if (nfaces < 0 || (nfaces > 0 && (!faces || !out)) || (pair != 1 && pair != 2) ||
        nfaces % pair)
        return HX_E_BADARG;
    if (nfaces == 0) return HX_OK;
    HX_TRY(hipSetDevice(c->device));
    FaceReq q{};
    q.faces = faces;
    q.nfaces = nfaces;
    q.pair = pair;
    q.accumulate = accumulate != 0;
    q.sx = scale_x;
    q.sy = scale_y;
    return launch_faces(c, solver, in, q, out, (hipStream_t)stream);
}

template <typename R>
static int fillfine_t(const hx_patch* co, const hx_patch* cn, double alpha, int ci0, int cj0,
                      const hx_patch* f, int whole, void* h, void* hu, void* hv, hipStream_t st,
                      const int* err = nullptr, const uint8_t* skip = nullptr,
                      int64_t spitch = 0)
{
    const int W = f->nx + 4, H = f->ny + 4;
    const int n = whole ? W * H : W * H - f->nx * f->ny;
    hipLaunchKernelGGL(fillfine_kernel<R>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       view_of<R>(co), view_of<R>(cn), (R)alpha, ci0, cj0, f->nx, f->ny, f->pitch,
                       whole, (const R*)f->z, (R*)h, (R*)hu, (R*)hv, err, skip, spitch);
    return hip_status(hipGetLastError());
}

extern "C" int hx_fill_fine_ghosts(hx_ctx* c, const hx_patch* co, const hx_patch* cn, double alpha,
                                   int ratio, int ci0, int cj0, const hx_patch* fine, int whole,
                                   void* h, void* hu, void* hv, void* stream)
{
    if (!c) return HX_E_STATE;
    int rc = check_patch(co, 0);
    if (!rc) rc = check_patch(cn, 0);
    if (rc) return rc;
    if (!fine || ratio != 2 || fine->ng < 2 || fine->nx % 2 || fine->ny % 2 || !h || !hu || !hv ||
        !fine->z || fine->dtype != co->dtype || cn->dtype != co->dtype ||
        fine->pitch < (int64_t)fine->nx + 4 || !(alpha >= 0.0 && alpha <= 1.0) ||
        co->nx != cn->nx || co->ny != cn->ny)
        return HX_E_BADARG;
    // the coarse stencil of the ghost ring (parents +- 1) must map into the
    // coarse level: parents lie within one coarse cell of the box
    const int ncx = fine->nx / 2, ncy = fine->ny / 2;
    if (ci0 < 0 || cj0 < 0 || ci0 + ncx > co->nx || cj0 + ncy > co->ny) return HX_E_BADARG;
    for (int d = 0; d < 2; ++d)
        if (co->bc_lo[d] == HX_BC_GHOST || co->bc_hi[d] == HX_BC_GHOST) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    return (co->dtype == HX_F64)
               ? fillfine_t<double>(co, cn, alpha, ci0, cj0, fine, whole != 0, h, hu, hv, st)
               : fillfine_t<float>(co, cn, alpha, ci0, cj0, fine, whole != 0, h, hu, hv, st);
}

template <typename R>
static int reflux_t(const hx_patch* crse, void* h, void* hu, void* hv, int ci0, int cj0, int ncx,
                    int ncy, const void* reg, hipStream_t st, const int* err = nullptr,
                    int orient = -1)
{
    const int nq = 2 * (ncx + ncy);
    hipLaunchKernelGGL(reflux_kernel<R>, dim3((nq + kBlock - 1) / kBlock), dim3(kBlock), 0, st, ci0,
                       cj0, ncx, ncy, crse->pitch, (R)(crse->dx * crse->dy), (const R*)reg, (R*)h,
                       (R*)hu, (R*)hv, err, orient, crse->nx, crse->ny);
    return hip_status(hipGetLastError());
}

extern "C" int hx_reflux(hx_ctx* c, const hx_patch* crse, void* h, void* hu, void* hv, int ci0,
                         int cj0, int ncx, int ncy, const void* reg, void* stream)
{
    if (!c) return HX_E_STATE;
    int rc = check_patch(crse, 0);
    if (rc) return rc;
    if (!h || !hu || !hv || !reg || ncx < 1 || ncy < 1 || ci0 < 1 || cj0 < 1 ||
        ci0 + ncx >= crse->nx || cj0 + ncy >= crse->ny)
        return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    return (crse->dtype == HX_F64) ? reflux_t<double>(crse, h, hu, hv, ci0, cj0, ncx, ncy, reg, st)
                                   : reflux_t<float>(crse, h, hu, hv, ci0, cj0, ncx, ncy, reg, st);
}

template <typename R>
static int avgdown_t(const hx_patch* fine, int ci0, int cj0, const hx_patch* crse, void* h,
                     void* hu, void* hv, hipStream_t st, const int* err = nullptr)
{
    const int ncx = fine->nx / 2, ncy = fine->ny / 2;
    hipLaunchKernelGGL(avgdown_state_kernel<R>, dim3((ncx * ncy + kBlock - 1) / kBlock), dim3(kBlock),
                       0, st, ci0, cj0, ncx, ncy, fine->pitch, (const R*)fine->h,
                       (const R*)fine->hu, (const R*)fine->hv, crse->pitch, (R*)h, (R*)hu, (R*)hv,
                       err);
    return hip_status(hipGetLastError());
}

extern "C" int hx_average_down_state(hx_ctx* c, const hx_patch* fine, int ratio, int ci0, int cj0,
                                     const hx_patch* crse, void* h, void* hu, void* hv,
                                     void* stream)
{
    if (!c) return HX_E_STATE;
    int rc = check_patch(fine, 0);
    if (!rc) rc = check_patch(crse, 0);
    if (rc) return rc;
    if (ratio != 2 || fine->nx % 2 || fine->ny % 2 || !h || !hu || !hv ||
        fine->dtype != crse->dtype)
        return HX_E_BADARG;
    const int ncx = fine->nx / 2, ncy = fine->ny / 2;
    if (ci0 < 0 || cj0 < 0 || ci0 + ncx > crse->nx || cj0 + ncy > crse->ny) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    return (crse->dtype == HX_F64) ? avgdown_t<double>(fine, ci0, cj0, crse, h, hu, hv, st)
                                   : avgdown_t<float>(fine, ci0, cj0, crse, h, hu, hv, st);
}

// ---------------------------------------------------------------------------
// Device-resident two-level run: the coupled step of the entry points above
// repeated without host synchronisation.  dt comes from the CFL reductions
// of both levels through the multi-level rule, on the device: the coarse
// advance reduces its output outside the box grown by one cell, a pass over
// the grown box reduces the cells reflux and average_down changed, and the
// last fine substep reduces the fine level.
// ---------------------------------------------------------------------------

// dt of a coupled step from the two levels' slots (include/solvers/BaseSolver.H:120-165
// at ratio 2; lockstep shares the finer dt, src/state/AmrMeshState_Physics.cpp:13-45),
// the same arithmetic as hx_level_dt + that rule on the host.  Clears the
// slots this step reduces into; on an estimate below 1e-12 on either level
// records the step in err and leaves dt alone (later kernels skip).
template <typename R>
__global__ void tl_dt_kernel(const CflSlot* rc, const CflSlot* rf, CflSlot* wc, CflSlot* wf,
                             double cfl, int subcycle, int step, double* dt, double* hist, int* err)
{
    if (*err) return;
    // the slots this step reduces into, one sub-slot a thread
    for (int q = threadIdx.x; q < kCflSubs; q += blockDim.x) {
        sub_clear(wc, q);
        sub_clear(wf, q);
    }
    const CflValue vc = slot_value_wave(rc, false), vf = slot_value_wave(rf, false);   // one wavefront
    if (threadIdx.x != 0) return;
    const R ec = slot_dt_est<R>(vc);
    const R ef = slot_dt_est<R>(vf);
    if ((double)ec < 1.e-12 || (double)ef < 1.e-12) {
        *err = step + 1;
        return;
    }
    double d0, d1;
    if (sizeof(R) == 8) {
        d0 = (double)ec * cfl;
        d1 = (double)ef * cfl;
    } else {
        d0 = (double)(ec * (R)cfl);
        d1 = (double)(ef * (R)cfl);
    }
    const double t0 = (d1 * 2.0 < d0) ? d1 * 2.0 : d0;
    const double a = subcycle ? t0 : t0 / 2.0, b = t0 / 2.0;
    dt[0] = a;
    dt[1] = b;
    hist[2 * step] = a;
    hist[2 * step + 1] = b;
}

// Coupled steps [s0, s1) (history and error indices relative to s0; the
// context's history holds kHistCap steps); fcur: the fine buffer holding the
// current fine state, carried across chunks.
template <typename R>
static int two_level_t(hx_ctx* c, int solver, const hx_patch* crse, void* const* calt,
                       const hx_patch* fine, void* const* falt, int ci0, int cj0, int subcycle,
                       double cfl, int s0, int s1, int& fcur, void* reg, TwoLevelWork* wk,
                       double* hist, hipStream_t st)
{
    const int ncx = fine->nx / 2, ncy = fine->ny / 2;
    CflSlot* slots = wk->slots;
    double* dt = wk->dt;
    int* err = &wk->err;
    int rc = HX_OK;
    if (s0 == 0) {
        hipLaunchKernelGGL(clear_slots_kernel, dim3(1), dim3(64), 0, st, slots, 4, err);
        rc = launch_cfl<R>(crse, slots + 0, st);
        if (!rc) rc = launch_cfl<R>(fine, slots + 2, st);
        if (rc) return rc;
    }
    void* cb[2][3] = {{(void*)crse->h, (void*)crse->hu, (void*)crse->hv},
                      {calt[0], calt[1], calt[2]}};
    void* fb[2][3] = {{(void*)fine->h, (void*)fine->hu, (void*)fine->hv},
                      {falt[0], falt[1], falt[2]}};
    StepArgs<R> ac{};
    ac.p = view_of<R>(crse);      // fill_geometry sizes the strips from p
    ac.opitch = crse->pitch;
    ac.cfl = cfl;
    ac.err = err;
    ac.dt_dev = dt;
    ac.sk_i0 = ci0 - 1; ac.sk_i1 = ci0 + ncx + 1;
    ac.sk_j0 = cj0 - 1; ac.sk_j1 = cj0 + ncy + 1;
    fill_geometry(ac, solver, eff_rows(c, solver, crse->dtype, crse->nx, crse->ny), crse->dx, crse->dy);
    StepArgs<R> af{};
    af.p = view_of<R>(fine);
    af.opitch = fine->pitch;
    af.cfl = cfl;
    af.err = err;
    af.dt_dev = dt + 1;
    fill_geometry(af, solver, eff_rows(c, solver, fine->dtype, fine->nx, fine->ny), fine->dx, fine->dy);
    FaceReq qc{};
    qc.nfaces = 2 * (ncx + ncy); qc.pair = 1; qc.accumulate = 0; qc.box_fine = 0;
    qc.ci0 = ci0; qc.cj0 = cj0; qc.ncx = ncx; qc.ncy = ncy;
    qc.sx = -crse->dy; qc.sy = -crse->dx; qc.dt = dt; qc.err = err;
    FaceReq qf = qc;
    qf.nfaces = 4 * (ncx + ncy); qf.pair = 2; qf.accumulate = 1; qf.box_fine = 1;
    qf.sx = fine->dy; qf.sy = fine->dx; qf.dt = dt + 1;
    const int nsub = subcycle ? 2 : 1;
    for (int s = s0; s < s1; ++s) {
        const int r = s & 1, w = r ^ 1;
        hipLaunchKernelGGL(tl_dt_kernel<R>, dim3(1), dim3(64), 0, st, slots + r, slots + 2 + r,
                           slots + w, slots + 2 + w, cfl, subcycle, s - s0, dt, hist, err);
        hx_patch cold = *crse, cnew = *crse;
        cold.h = cb[r][0]; cold.hu = cb[r][1]; cold.hv = cb[r][2];
        cnew.h = cb[w][0]; cnew.hu = cb[w][1]; cnew.hv = cb[w][2];
        // coarse register from the step's input, then the coarse advance
        rc = launch_faces(c, solver, &cold, qc, reg, st);
        if (rc) return rc;
        ac.p = view_of<R>(&cold);
        ac.ho = (R*)cb[w][0]; ac.huo = (R*)cb[w][1]; ac.hvo = (R*)cb[w][2];
        ac.acc = slots + w;
        ac.step = s;
        rc = launch(c, solver, crse->dtype, ac, st);
        if (rc) return rc;
        for (int k = 0; k < nsub; ++k) {
            hx_patch fp = *fine;
            fp.h = fb[fcur][0]; fp.hu = fb[fcur][1]; fp.hv = fb[fcur][2];
            rc = fillfine_t<R>(&cold, &cnew, k / 2.0, ci0, cj0, &fp, 0, fb[fcur][0], fb[fcur][1],
                               fb[fcur][2], st, err);
            if (!rc) rc = launch_faces(c, solver, &fp, qf, reg, st);
            if (rc) return rc;
            af.p = view_of<R>(&fp);
            af.ho = (R*)fb[fcur ^ 1][0]; af.huo = (R*)fb[fcur ^ 1][1]; af.hvo = (R*)fb[fcur ^ 1][2];
            af.acc = (k == nsub - 1) ? slots + 2 + w : nullptr;
            af.step = s;
            rc = launch(c, solver, fine->dtype, af, st);
            if (rc) return rc;
            fcur ^= 1;
        }
        rc = reflux_t<R>(&cnew, cb[w][0], cb[w][1], cb[w][2], ci0, cj0, ncx, ncy, reg, st, err);
        hx_patch ffin = *fine;
        ffin.h = fb[fcur][0]; ffin.hu = fb[fcur][1]; ffin.hv = fb[fcur][2];
        if (!rc) rc = avgdown_t<R>(&ffin, ci0, cj0, &cnew, cb[w][0], cb[w][1], cb[w][2], st, err);
        if (rc) return rc;
        // CFL of the grown box (covered cells, the reflux ring and its corners)
        hx_patch gb = cnew;
        const int64_t off = (int64_t)(cj0 - 1) * crse->pitch + (ci0 - 1);
        const size_t es = (crse->dtype == HX_F64) ? 8 : 4;
        gb.h = (const char*)cnew.h + off * es;
        gb.hu = (const char*)cnew.hu + off * es;
        gb.hv = (const char*)cnew.hv + off * es;
        gb.z = (const char*)crse->z + off * es;
        gb.n = (const char*)crse->n + off * es;
        gb.nx = ncx + 2;
        gb.ny = ncy + 2;
        rc = launch_cfl<R>(&gb, slots + w, st);
        if (rc) return rc;
    }
    return HX_OK;
}

// Arguments of a coupled two-level run (hx_two_level_run, hx_two_level_enqueue).
static int two_level_check(const hx_ctx* c, int solver, const hx_patch* crse, void* ch, void* chu,
                           void* chv, const hx_patch* fine, void* fh, void* fhu, void* fhv,
                           int ci0, int cj0, double cfl, int nsteps, void* reg)
{
    if (!c) return HX_E_STATE;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    int rc = check_patch(crse, 0);
    if (!rc) rc = check_patch(fine, 2);
    if (rc) return rc;
    for (int d = 0; d < 2; ++d) {
        if (crse->bc_lo[d] == HX_BC_GHOST || crse->bc_hi[d] == HX_BC_GHOST) return HX_E_BADARG;
        if (fine->bc_lo[d] != HX_BC_GHOST || fine->bc_hi[d] != HX_BC_GHOST) return HX_E_BADARG;
    }
    if (crse->eta || fine->eta || !ch || !chu || !chv || !fh || !fhu || !fhv || !reg ||
        nsteps < 0 || !(cfl > 0.0) || fine->dtype != crse->dtype || fine->ng < 2 ||
        fine->nx % 2 || fine->ny % 2 || fine->nx < 2 || fine->ny < 2 ||
        fine->dx != 0.5 * crse->dx || fine->dy != 0.5 * crse->dy)
        return HX_E_BADARG;
    const int ncx = fine->nx / 2, ncy = fine->ny / 2;
    if (ci0 < 1 || cj0 < 1 || ci0 + ncx > crse->nx - 1 || cj0 + ncy > crse->ny - 1)
        return HX_E_BADARG;
    return HX_OK;
}

extern "C" int hx_two_level_run(hx_ctx* c, int solver, const hx_patch* crse, void* ch, void* chu,
                                void* chv, const hx_patch* fine, void* fh, void* fhu, void* fhv,
                                int ci0, int cj0, int subcycle, double cfl, int nsteps, void* reg,
                                double* dt_hist, int32_t* steps_done, void* stream)
{
    int rc = two_level_check(c, solver, crse, ch, chu, chv, fine, fh, fhu, fhv, ci0, cj0, cfl,
                             nsteps, reg);
    if (rc) return rc;
    if (steps_done) *steps_done = 0;
    if (nsteps == 0) return HX_OK;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    void* calt[3] = {ch, chu, chv};
    void* falt[3] = {fh, fhu, fhv};
    int fcur = 0;
    for (int s0 = 0; s0 < nsteps; s0 += c->hist_cap) {
        const int s1 = std::min(nsteps, s0 + c->hist_cap);
        rc = (crse->dtype == HX_F64)
                 ? two_level_t<double>(c, solver, crse, calt, fine, falt, ci0, cj0, subcycle != 0,
                                       cfl, s0, s1, fcur, reg, c->tlw, c->hist, st)
                 : two_level_t<float>(c, solver, crse, calt, fine, falt, ci0, cj0, subcycle != 0,
                                      cfl, s0, s1, fcur, reg, c->tlw, c->hist, st);
        if (rc) return rc;
        // read back through the context's pinned staging: [error flag][dt pairs]
        int* perr = (int*)c->hpin;
        double* pdt = c->hpin + 1;
        HX_TRY(hipMemcpyAsync(perr, &c->tlw->err, sizeof(int), hipMemcpyDeviceToHost, st));
        if (dt_hist)
            HX_TRY(hipMemcpyAsync(pdt, c->hist, sizeof(double) * 2 * (s1 - s0),
                                  hipMemcpyDeviceToHost, st));
        HX_TRY(hipStreamSynchronize(st));
        const int err = *perr;
        if (dt_hist) std::memcpy(dt_hist + 2 * (size_t)s0, pdt, sizeof(double) * 2 * (s1 - s0));
        if (err) {
            if (steps_done) *steps_done = s0 + err - 1;
            return HX_E_DT_COLLAPSE;
        }
    }
    if (steps_done) *steps_done = nsteps;
    return HX_OK;
}

__global__ void tl_status_kernel(const int* err, int32_t* out)
{
    if (threadIdx.x == 0) *out = *err;
}

extern "C" int hx_two_level_enqueue(hx_ctx* c, int solver, const hx_patch* crse, void* ch,
                                    void* chu, void* chv, const hx_patch* fine, void* fh,
                                    void* fhu, void* fhv, int ci0, int cj0, int subcycle,
                                    double cfl, int nsteps, void* reg, double* dt_hist_dev,
                                    int32_t* status_dev, void* stream)
{
    int rc = two_level_check(c, solver, crse, ch, chu, chv, fine, fh, fhu, fhv, ci0, cj0, cfl,
                             nsteps, reg);
    if (rc) return rc;
    if (nsteps < 1 || nsteps > c->hist_cap) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    void* calt[3] = {ch, chu, chv};
    void* falt[3] = {fh, fhu, fhv};
    int fcur = 0;
    double* hist = dt_hist_dev ? dt_hist_dev : c->hist;
    rc = (crse->dtype == HX_F64)
             ? two_level_t<double>(c, solver, crse, calt, fine, falt, ci0, cj0, subcycle != 0, cfl,
                                   0, nsteps, fcur, reg, c->tlw, hist, st)
             : two_level_t<float>(c, solver, crse, calt, fine, falt, ci0, cj0, subcycle != 0, cfl,
                                  0, nsteps, fcur, reg, c->tlw, hist, st);
    if (rc) return rc;
    if (status_dev) {
        hipLaunchKernelGGL(tl_status_kernel, dim3(1), dim3(64), 0, st, &c->tlw->err, status_dev);
        HX_TRY(hipGetLastError());
    }
    return HX_OK;
}

// ---------------------------------------------------------------------------
// Entry points of the multi-box hierarchy (dense levels, box lists)
// ---------------------------------------------------------------------------
extern "C" int hx_fill_fine_region(hx_ctx* c, const hx_patch* co, const hx_patch* cn, double alpha,
                                   int ratio, int ci0, int cj0, const hx_patch* fine, int whole,
                                   const uint8_t* skip, int64_t skip_pitch, void* h, void* hu,
                                   void* hv, void* stream)
{
    if (!c) return HX_E_STATE;
    int rc = check_patch(co, 0);
    if (!rc) rc = check_patch(cn, 0);
    if (rc) return rc;
    if (!fine || ratio != 2 || fine->ng < 2 || fine->nx % 2 || fine->ny % 2 || !h || !hu || !hv ||
        !fine->z || fine->dtype != co->dtype || cn->dtype != co->dtype ||
        fine->pitch < (int64_t)fine->nx + 4 || !(alpha >= 0.0 && alpha <= 1.0) ||
        co->nx != cn->nx || co->ny != cn->ny || (skip && skip_pitch < (int64_t)fine->nx + 4))
        return HX_E_BADARG;
    const int ncx = fine->nx / 2, ncy = fine->ny / 2;
    if (ci0 < 0 || cj0 < 0 || ci0 + ncx > co->nx || cj0 + ncy > co->ny) return HX_E_BADARG;
    for (int d = 0; d < 2; ++d)
        if (co->bc_lo[d] == HX_BC_GHOST || co->bc_hi[d] == HX_BC_GHOST) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    return (co->dtype == HX_F64)
               ? fillfine_t<double>(co, cn, alpha, ci0, cj0, fine, whole != 0, h, hu, hv, st,
                                    nullptr, skip, skip_pitch)
               : fillfine_t<float>(co, cn, alpha, ci0, cj0, fine, whole != 0, h, hu, hv, st,
                                   nullptr, skip, skip_pitch);
}

extern "C" int hx_reflux_side(hx_ctx* c, const hx_patch* crse, void* h, void* hu, void* hv,
                              int ci0, int cj0, int ncx, int ncy, const void* reg, int orient,
                              void* stream)
{
    if (!c) return HX_E_STATE;
    int rc = check_patch(crse, 0);
    if (rc) return rc;
    if (!h || !hu || !hv || !reg || ncx < 1 || ncy < 1 || ci0 < 0 || cj0 < 0 ||
        ci0 + ncx > crse->nx || cj0 + ncy > crse->ny || orient < -1 || orient > 3)
        return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    return (crse->dtype == HX_F64)
               ? reflux_t<double>(crse, h, hu, hv, ci0, cj0, ncx, ncy, reg, st, nullptr, orient)
               : reflux_t<float>(crse, h, hu, hv, ci0, cj0, ncx, ncy, reg, st, nullptr, orient);
}

extern "C" int hx_boxes_dt(hx_ctx* c, int solver, int nbox, const hx_patch* boxes, double cfl,
                           double* dt_out, double* max_vel_out, void* stream)
{
    if (!c) return HX_E_STATE;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    if (nbox < 1 || !boxes || !dt_out || !(cfl > 0.0)) return HX_E_BADARG;
    for (int b = 0; b < nbox; ++b) {
        const int rc = check_patch(&boxes[b], 0);
        if (rc) return rc;
        if (boxes[b].dtype != boxes[0].dtype) return HX_E_BADARG;
    }
    HX_TRY(hipSetDevice(c->device));
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(clear_slots_kernel, dim3(1), dim3(64), 0, st, c->slots + 3, 1, (int*)nullptr);
    for (int b = 0; b < nbox; ++b) {
        const int rc = (boxes[b].dtype == HX_F64) ? launch_cfl<double>(&boxes[b], c->slots + 3, st)
                                                  : launch_cfl<float>(&boxes[b], c->slots + 3, st);
        if (rc) return rc;
    }
    CflSlot slot;
    HX_TRY(hipMemcpyAsync(c->hpin, c->slots + 3, sizeof(slot), hipMemcpyDeviceToHost, st));
    HX_TRY(hipStreamSynchronize(st));
    std::memcpy(&slot, c->hpin, sizeof(slot));
    const CflValue sl = slot_value(&slot);
    double vmax;
    std::memcpy(&vmax, &sl.vmax, sizeof(vmax));
    double est;
    if (boxes[0].dtype == HX_F64) {
        est = slot_dt_est<double>(sl);
        *dt_out = est * cfl;
    } else {
        const float e = slot_dt_est<float>(sl);
        est = e;
        *dt_out = (double)(e * (float)cfl);
    }
    if (max_vel_out) *max_vel_out = (sl.flags ? vmax : -1.7976931348623157e308);
    return (est < 1.e-12) ? HX_E_DT_COLLAPSE : HX_OK;
}

// ---------------------------------------------------------------------------
// Native AMR run: the recursive step of AmrMeshState::TimeStepWithSubcycling
// (src/state/AmrMeshState_Physics.cpp:50-147) with ComputeDt (:13-45) over
// levels held whole in device memory, launched from C++.  The host is called
// back only to regrid.
// ---------------------------------------------------------------------------
namespace {

struct AmrRun {
    hx_ctx* c;
    int solver, dtype;
    hx_amr_level* lv;
    int* nlev;
    int max_level, subcycle, reflux, regrid_int;
    int32_t bclo[2], bchi[2];
    hx_amr_regrid_fn regrid;
    void* user;
    double dt[HX_AMR_MAX_LEVELS];
    CflSlot* slots;     // device, one per level (the context's)
    hipStream_t st;
    // patch tables of every level, concatenated (host copy and the
    // context's device buffer), and each level's launch sizes
    std::vector<AmrPatch> tab;
    AmrPatch* dtab;
    int toff[HX_AMR_MAX_LEVELS], tn[HX_AMR_MAX_LEVELS];
    int strips[HX_AMR_MAX_LEVELS], cfl_blocks[HX_AMR_MAX_LEVELS];
    int fill_blocks[HX_AMR_MAX_LEVELS], face_blocks[HX_AMR_MAX_LEVELS];
    int avg_blocks[HX_AMR_MAX_LEVELS];
    bool upload_pending;   // a table upload from the pinned staging may be in flight
};

// A level's arrays for the patch-table kernels: u_in and u_out the state
// buffers bi and bo.
template <typename R>
AmrLevelArgs<R> amr_args(const AmrRun& r, int l, int bi, int bo)
{
    const hx_amr_level& L = r.lv[l];
    AmrLevelArgs<R> a;
    a.pitch = L.nx + 4;
    a.plane = (int64_t)(L.ny + 4) * a.pitch;
    a.u_in = (const R*)L.u[bi];
    a.u_out = (R*)L.u[bo];
    a.terrain = (const R*)L.terrain;
    a.mask = L.mask;
    a.nx = L.nx; a.ny = L.ny;
    a.bclo0 = r.bclo[0]; a.bclo1 = r.bclo[1]; a.bchi0 = r.bchi[0]; a.bchi1 = r.bchi[1];
    a.dx = (R)L.dx; a.dy = (R)L.dy;
    a.rdx = (R)(1.0 / L.dx); a.rdy = (R)(1.0 / L.dy);
    a.dx_pow2 = is_pow2(L.dx) && is_pow2(L.dy) ? 1 : 0;
    return a;
}

int amr_reg_offset(const hx_amr_level& F, int k);

// Patch tables of every level from the hosts' box lists (at the start of a
// run and after each regrid), uploaded to the context's table buffer, which
// grows when a hierarchy has more boxes than it holds.
template <typename R>
int amr_tables(AmrRun& r)
{
    const int cols = (r.solver == HX_SOLVER_ROE) ? kRoeCols : kHydroCols;
    r.tab.clear();
    for (int l = 0; l < *r.nlev; ++l) {
        const hx_amr_level& L = r.lv[l];
        const int64_t pitch = L.nx + 4;
        r.toff[l] = (int)r.tab.size();
        r.tn[l] = L.npatch;
        int strips = 0, cfl = 0, fill = 0, face = 0, avg = 0;
        for (int k = 0; k < L.npatch; ++k) {
            const int32_t* b = L.patches + 4 * k;
            AmrPatch P;
            P.i0 = b[0]; P.j0 = b[1]; P.w = b[2]; P.h = b[3];
            P.rows = clamp_strip_rows<R>(eff_rows(r.c, r.solver, r.dtype, P.w, P.h), pitch, pitch);
            P.nsx = (P.w + cols - 1) / cols;
            P.strip0 = strips;
            strips += P.nsx * ((P.h + P.rows - 1) / P.rows);
            P.cflx = (P.w + kBlock - 1) / kBlock;
            P.cfl0 = cfl;
            cfl += P.cflx * ((P.h + kCflRows - 1) / kCflRows);
            const int ring = (P.w + 4) * (P.h + 4) - P.w * P.h;
            P.fill0 = fill;
            fill += (ring + kBlock - 1) / kBlock;
            P.face0 = face;
            face += (2 * (P.w / 2 + P.h / 2) + kBlock - 1) / kBlock;
            P.avg0 = avg;
            avg += ((P.w / 2) * (P.h / 2) + kBlock - 1) / kBlock;
            P.reg = amr_reg_offset(L, k);
            r.tab.push_back(P);
        }
        r.strips[l] = strips; r.cfl_blocks[l] = cfl; r.fill_blocks[l] = fill;
        r.face_blocks[l] = face; r.avg_blocks[l] = avg;
    }
    hx_ctx* c = r.c;
    const int n = (int)r.tab.size();
    const size_t bytes = sizeof(AmrPatch) * n;
    r.dtab = c->amr_tab;
    // the device already holds this table (a regrid that kept the grids)
    if (n == c->amr_n && (n == 0 || std::memcmp(c->amr_host, r.tab.data(), bytes) == 0))
        return HX_OK;
    if (n > c->amr_cap) {
        const int cap = std::max(n, 2 * c->amr_cap + 256);
        // kernels queued earlier may still read the old table
        HX_TRY(hipStreamSynchronize(r.st));
        if (c->amr_tab) HX_TRY(hipFree(c->amr_tab));
        if (c->amr_host) HX_TRY(hipHostFree(c->amr_host));
        c->amr_tab = nullptr;
        c->amr_host = nullptr;
        c->amr_cap = 0;
        c->amr_n = -1;
        if (hipMalloc((void**)&c->amr_tab, sizeof(AmrPatch) * cap) != hipSuccess ||
            hipHostMalloc((void**)&c->amr_host, sizeof(AmrPatch) * cap) != hipSuccess)
            return HX_E_NOMEM;
        c->amr_cap = cap;
        r.dtab = c->amr_tab;
    } else if (r.upload_pending) {
        // the pinned staging copy of an earlier upload may still be in flight
        HX_TRY(hipStreamSynchronize(r.st));
    }
    std::memcpy(c->amr_host, r.tab.data(), bytes);
    c->amr_n = n;
    if (n) {
        // stream-ordered: launches queued before read the old table, those after the new
        HX_TRY(hipMemcpyAsync(r.dtab, c->amr_host, bytes, hipMemcpyHostToDevice, r.st));
        r.upload_pending = true;
    }
    return HX_OK;
}

bool almost_equal(double x, double y)
{
    const double m = std::max(std::max(std::fabs(x), std::fabs(y)), 1e-300);
    return std::fabs(x - y) <= 1e-12 * m;
}

template <typename R>
int amr_fill(AmrRun& r, int l, double t)
{
    hx_amr_level &C = r.lv[l - 1], &F = r.lv[l];
    int bo, bn;
    double alpha = 0.0;
    if (almost_equal(t, C.t_new)) { bo = bn = C.cur; }
    else if (almost_equal(t, C.t_old)) { bo = bn = C.cur ^ 1; }
    else { bo = C.cur ^ 1; bn = C.cur; alpha = (t - C.t_old) / (C.t_new - C.t_old); }
    if (r.fill_blocks[l] == 0) return HX_OK;
    const AmrLevelArgs<R> ca = amr_args<R>(r, l - 1, C.cur, C.cur);
    const AmrLevelArgs<R> fa = amr_args<R>(r, l, F.cur, F.cur);
    hipLaunchKernelGGL(amr_fill_kernel<R>, dim3(r.fill_blocks[l]), dim3(kBlock), 0, r.st, ca,
                       (const R*)C.u[bo], (const R*)C.u[bn], (R)alpha, fa, r.dtab + r.toff[l],
                       r.tn[l]);
    return hip_status(hipGetLastError());
}

int amr_reg_offset(const hx_amr_level& F, int k)
{
    int off = 0;
    for (int q = 0; q < k; ++q) off += 6 * (F.patches[4 * q + 2] / 2 + F.patches[4 * q + 3] / 2);
    return off;
}

template <typename R>
int amr_advance(AmrRun& r, int l, double t, double dt)
{
    hx_amr_level& L = r.lv[l];
    const bool fast = (r.c->arith == HX_ARITH_FAST);
    int rc;
    if (l > 0 && (rc = amr_fill<R>(r, l, t))) return rc;
    if (r.reflux && l < *r.nlev - 1) {            // CrseInit for every patch of level l + 1
        hx_amr_level& F = r.lv[l + 1];
        const AmrLevelArgs<R> la = amr_args<R>(r, l, L.cur, L.cur);
        rc = fast ? fast::launch_amr_faces<R>(r.solver, la, r.dtab + r.toff[l + 1], r.tn[l + 1],
                                              r.face_blocks[l + 1], 0, -dt * L.dy, -dt * L.dx,
                                              (R*)F.regs, r.st)
                  : exact::launch_amr_faces<R>(r.solver, la, r.dtab + r.toff[l + 1], r.tn[l + 1],
                                               r.face_blocks[l + 1], 0, -dt * L.dy, -dt * L.dx,
                                               (R*)F.regs, r.st);
        if (rc) return HX_E_HIP;
    }
    if (r.reflux && l > 0) {                      // FineAdd of this level's patches
        const AmrLevelArgs<R> la = amr_args<R>(r, l, L.cur, L.cur);
        rc = fast ? fast::launch_amr_faces<R>(r.solver, la, r.dtab + r.toff[l], r.tn[l],
                                              r.face_blocks[l], 1, dt * L.dy, dt * L.dx,
                                              (R*)L.regs, r.st)
                  : exact::launch_amr_faces<R>(r.solver, la, r.dtab + r.toff[l], r.tn[l],
                                               r.face_blocks[l], 1, dt * L.dy, dt * L.dx,
                                               (R*)L.regs, r.st);
        if (rc) return HX_E_HIP;
    }
    L.cur ^= 1;
    const AmrLevelArgs<R> la = amr_args<R>(r, l, L.cur ^ 1, L.cur);
    const int occ = eff_occ(r.c, r.solver, r.dtype);
    rc = fast ? fast::launch_amr_advance<R>(r.solver, occ, la, r.dtab + r.toff[l], r.tn[l],
                                            r.strips[l], (R)dt, r.st)
              : exact::launch_amr_advance<R>(r.solver, occ, la, r.dtab + r.toff[l], r.tn[l],
                                             r.strips[l], (R)dt, r.st);
    return rc ? HX_E_HIP : HX_OK;
}

template <typename R>
int amr_reflux_avgdown(AmrRun& r, int l)
{
    hx_amr_level &L = r.lv[l], &F = r.lv[l + 1];
    const AmrLevelArgs<R> ca = amr_args<R>(r, l, L.cur, L.cur);
    const AmrPatch* ft = r.dtab + r.toff[l + 1];
    const int fn = r.tn[l + 1];
    if (r.reflux && r.face_blocks[l + 1] > 0)
        for (int orient = 0; orient < 4; ++orient) {   // AMReX's side order, all boxes per side
            hipLaunchKernelGGL(amr_reflux_kernel<R>, dim3(r.face_blocks[l + 1]), dim3(kBlock), 0,
                               r.st, ca, ft, fn, (const R*)F.regs, orient, (R)(L.dx * L.dy));
            const int rc = hip_status(hipGetLastError());
            if (rc) return rc;
        }
    if (r.avg_blocks[l + 1] > 0) {
        const AmrLevelArgs<R> fa = amr_args<R>(r, l + 1, F.cur, F.cur);
        hipLaunchKernelGGL(amr_avgdown_kernel<R>, dim3(r.avg_blocks[l + 1]), dim3(kBlock), 0, r.st,
                           fa, ca, ft, fn);
        return hip_status(hipGetLastError());
    }
    return HX_OK;
}

template <typename R>
int amr_time_step(AmrRun& r, int l, double time)
{
    int rc;
    hx_amr_level* L = &r.lv[l];
    if (r.regrid_int > 0 && l < r.max_level && L->istep > L->last_regrid &&
        L->istep % r.regrid_int == 0 && r.regrid) {
        const int old_n = *r.nlev;
        if ((rc = r.regrid(r.user, l, time))) return rc;
        for (int k = l; k < *r.nlev; ++k) r.lv[k].last_regrid = r.lv[k].istep;
        for (int k = old_n; k < *r.nlev; ++k) r.dt[k] = r.dt[k - 1] / (r.subcycle ? 2.0 : 1.0);
        if ((rc = amr_tables<R>(r))) return rc;
        L = &r.lv[l];
    }
    L->t_old = L->t_new;
    L->t_new += r.dt[l];
    if ((rc = amr_advance<R>(r, l, L->t_old, r.dt[l]))) return rc;
    L->istep += 1;
    if (l < *r.nlev - 1) {
        const int nsub = r.subcycle ? 2 : 1;
        for (int i = 1; i <= nsub; ++i)
            if ((rc = amr_time_step<R>(r, l + 1, time + (i - 1) * r.dt[l + 1]))) return rc;
        if ((rc = amr_reflux_avgdown<R>(r, l))) return rc;
    }
    return HX_OK;
}

// BaseSolver::compute_dt at ratio 2 and the lockstep rule: every level's
// patches reduce into that level's slot, one synchronisation.
template <typename R>
int amr_compute_dt(AmrRun& r, double cfl, int* collapsed)
{
    const int n = *r.nlev;
    hipLaunchKernelGGL(clear_slots_kernel, dim3(1), dim3(64), 0, r.st, r.slots, n, (int*)nullptr);
    for (int l = 0; l < n; ++l) {
        const hx_amr_level& L = r.lv[l];
        if (r.cfl_blocks[l] == 0) continue;
        const AmrLevelArgs<R> la = amr_args<R>(r, l, L.cur, L.cur);
        hipLaunchKernelGGL(amr_cfl_kernel<R>, dim3(r.cfl_blocks[l]), dim3(kBlock), 0, r.st, la,
                           r.dtab + r.toff[l], r.tn[l], la.dx_pow2, r.slots + l);
        const int rc = hip_status(hipGetLastError());
        if (rc) return rc;
    }
    // read back through the context's pinned staging (HX_AMR_MAX_LEVELS slots
    // fit in it many times over)
    CflValue s[HX_AMR_MAX_LEVELS];
    HX_TRY(hipMemcpyAsync(r.c->hpin, r.slots, sizeof(CflSlot) * n, hipMemcpyDeviceToHost, r.st));
    HX_TRY(hipStreamSynchronize(r.st));
    for (int l = 0; l < n; ++l) s[l] = slot_value((const CflSlot*)r.c->hpin + l);
    r.upload_pending = false;
    double tmp[HX_AMR_MAX_LEVELS];
    *collapsed = 0;
    for (int l = 0; l < n; ++l) {
        double est;
        if (sizeof(R) == 8) {
            est = slot_dt_est<double>(s[l].amax, s[l].flags);
            tmp[l] = est * cfl;
        } else {
            const float e = slot_dt_est<float>(s[l].amax, s[l].flags);
            est = e;
            tmp[l] = (double)(e * (float)cfl);
        }
        if (est < 1.e-12) *collapsed = 1;
    }
    double dt0 = tmp[0], factor = 1.0;
    for (int l = 1; l < n; ++l) {
        factor *= 2.0;
        dt0 = std::min(dt0, tmp[l] * factor);
    }
    r.dt[0] = dt0;
    for (int l = 1; l < n; ++l) r.dt[l] = r.dt[l - 1] / 2.0;
    if (!r.subcycle) {
        double m = r.dt[0];
        for (int l = 1; l < n; ++l) m = std::min(m, r.dt[l]);
        for (int l = 0; l < n; ++l) r.dt[l] = m;
    }
    return HX_OK;
}

}  // namespace

extern "C" int hx_amr_run(hx_ctx* c, int solver, int dtype, hx_amr_level* levels, int* nlevels,
                          int max_level, const int32_t* bc_lo, const int32_t* bc_hi, int subcycle,
                          int do_reflux, int regrid_int, hx_amr_regrid_fn regrid, void* user,
                          double cfl, double t_end, int max_steps, double* time,
                          double* dt_hist, int32_t* steps_done, void* stream)
{
    if (!c) return HX_E_STATE;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    if (!levels || !nlevels || *nlevels < 1 || *nlevels > max_level + 1 || !bc_lo || !bc_hi ||
        max_level + 1 > HX_AMR_MAX_LEVELS || !time || !(cfl > 0.0) || max_steps < 0 ||
        (dtype != HX_F64 && dtype != HX_F32))
        return HX_E_BADARG;
    for (int d = 0; d < 2; ++d)
        if ((bc_lo[d] != 0 && bc_lo[d] != 1) || (bc_hi[d] != 0 && bc_hi[d] != 1)) return HX_E_BADARG;
    if (steps_done) *steps_done = 0;
    HX_TRY(hipSetDevice(c->device));
    AmrRun r{};
    r.c = c; r.solver = solver; r.dtype = dtype; r.lv = levels; r.nlev = nlevels;
    r.max_level = max_level; r.subcycle = subcycle != 0; r.reflux = do_reflux != 0;
    r.regrid_int = regrid_int; r.regrid = regrid; r.user = user;
    r.bclo[0] = bc_lo[0]; r.bclo[1] = bc_lo[1]; r.bchi[0] = bc_hi[0]; r.bchi[1] = bc_hi[1];
    r.st = (hipStream_t)stream;
    r.slots = c->amr_slots;
    int rc = (dtype == HX_F64) ? amr_tables<double>(r) : amr_tables<float>(r);
    if (rc) return rc;
    int k = 0;
    while (*time < t_end && k < max_steps) {
        int collapsed = 0;
        rc = (dtype == HX_F64) ? amr_compute_dt<double>(r, cfl, &collapsed)
                               : amr_compute_dt<float>(r, cfl, &collapsed);
        if (rc) break;
        if (collapsed) { rc = HX_E_DT_COLLAPSE; break; }
        if (dt_hist) dt_hist[k] = r.dt[0];
        rc = (dtype == HX_F64) ? amr_time_step<double>(r, 0, *time)
                               : amr_time_step<float>(r, 0, *time);
        if (rc) break;
        *time = r.lv[0].t_new;
        ++k;
    }
    if (hipStreamSynchronize(r.st) != hipSuccess && !rc) rc = HX_E_HIP;
    if (steps_done) *steps_done = k;
    return rc;
}

// ---------------------------------------------------------------------------
// Regrid tag blocks: block (bi, bj) of b x b cells is set when one of its
// valid cells lies within nbuf cells (max norm) of a tagged valid cell --
// the tags grown by n_error_buf, kept on the level's valid cells and
// coarsened by the blocking factor in one pass, so only the block array
// travels to the host.
// ---------------------------------------------------------------------------
namespace {
__global__ __launch_bounds__(256) void tag_blocks_kernel(const uint8_t* __restrict__ tags,
                                                         int64_t tpitch,
                                                         const uint8_t* __restrict__ valid,
                                                         int64_t vpitch, int nx, int ny, int nbuf,
                                                         int b, uint8_t* __restrict__ blocks)
{
    const int nbx = nx / b, nby = ny / b;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nbx * nby) return;
    const int bi = q % nbx, bj = q / nbx;
    uint8_t hit = 0;
    for (int j = bj * b; j < bj * b + b && !hit; ++j)
        for (int i = bi * b; i < bi * b + b && !hit; ++i) {
            if (!valid[(int64_t)j * vpitch + i]) continue;
            for (int dj = -nbuf; dj <= nbuf && !hit; ++dj) {
                const int y = j + dj;
                if (y < 0 || y >= ny) continue;
                for (int di = -nbuf; di <= nbuf; ++di) {
                    const int x = i + di;
                    if (x < 0 || x >= nx) continue;
                    if (tags[(int64_t)y * tpitch + x] && valid[(int64_t)y * vpitch + x]) {
                        hit = 1;
                        break;
                    }
                }
            }
        }
    blocks[q] = hit;
}
}  // namespace

extern "C" int hx_tag_blocks(hx_ctx* c, const uint8_t* tags, int64_t tag_pitch,
                             const uint8_t* valid, int64_t valid_pitch, int nx, int ny, int nbuf,
                             int b, uint8_t* blocks, void* stream)
{
    if (!c) return HX_E_STATE;
    if (!tags || !valid || !blocks || nx < 1 || ny < 1 || nbuf < 0 || b < 1 || nx % b || ny % b ||
        tag_pitch < nx || valid_pitch < nx)
        return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    const int n = (nx / b) * (ny / b);
    hipLaunchKernelGGL(tag_blocks_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       tags, tag_pitch, valid, valid_pitch, nx, ny, nbuf, b, blocks);
    return hip_status(hipGetLastError());
}
