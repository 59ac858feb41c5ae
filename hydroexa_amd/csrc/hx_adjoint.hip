// hx_adjoint.hip -- adjoint (reverse-mode) of the per-level step with
// checkpointing, FP64.  No reference counterpart: HydroEXA has no
// derivatives.  The forward trajectory is the library's own fused advance;
// the adjoint linearises each step around the state the forward produced,
// with the dt schedule of that forward held fixed.
//
// One adjoint step, lambda^{n+1} -> lambda^n (plus the roughness and bed
// gradients accumulated), is deterministic (no atomics):
//   post    (hydrostatic only): per cell, the reverse of Manning friction
//           and the dry clamp at the step's state after the flux update,
//           which the segment's forward recomputation stored (StepArgs::uno)
//           -> lambda_u (adjoint of the pre-friction state);
//   tile    per tile of cells: the vector-Jacobian product of every face
//           whose stencil touches the tile (2 cells for plain Roe, 4 for the
//           hydrostatic face), summed in LDS, then per cell the gather over
//           its own position and the ghosts filled from it, with the
//           ghost-fill and rebuild chain rules (eta = h + z, h_rebuilt =
//           max(0, eta - z), mirror / edge sources, odd momenta).
// Compiled with -ffp-contract=fast: derivatives need relative accuracy,
// not the forward's bit-exact rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "../../include/hydroexa.h"
#include "hx_adjoint.hpp"
#include "hx_host.hpp"
#include "hx_types.hpp"

namespace hx {
namespace adj {

constexpr int kThreads = 256;

struct Geo {
    int nx, ny, bl0, bl1, bh0, bh1;
};

// One state (dense planes, row pitch nx) with its terrain.
struct State {
    const double* h;
    const double* hu;
    const double* hv;
    const double* z;
    const double* n;
};

// Physical-BC source of a padded index (include/boundaries/BCFill.H:49-82,
// terrain foextrap per src/state/AmrMeshState_Init.cpp:62-70): src for the
// state, edge for the terrain, flip for odd (normal) momenta.
__device__ __forceinline__ void map1(int i, int n, int blo, int bhi, int& src, int& edge,
                                     bool& flip)
{
    src = i;
    edge = i;
    flip = false;
    if (i < 0) {
        edge = 0;
        flip = (blo == 0);
        src = flip ? (-1 - i) : 0;
    } else if (i >= n) {
        edge = n - 1;
        flip = (bhi == 0);
        src = flip ? (2 * n - 1 - i) : n - 1;
    }
}

// Padded cell as the forward step sees it; x = eta - z before the max.
struct PC {
    double he, hu, hv, z, n, x;
};

__device__ __forceinline__ PC pcell(const State& s, const Geo& g, int i, int j)
{
    int si, ei, sj, ej;
    bool fx, fy;
    map1(i, g.nx, g.bl0, g.bh0, si, ei, fx);
    map1(j, g.ny, g.bl1, g.bh1, sj, ej, fy);
    const int64_t a = (int64_t)sj * g.nx + si, b = (int64_t)ej * g.nx + ei;
    PC c;
    c.z = s.z[b];
    c.n = s.n[b];
    const double eta = s.h[a] + s.z[a];
    c.x = eta - c.z;
    c.he = (0.0 < c.x) ? c.x : 0.0;
    const double qx = s.hu[a], qy = s.hv[a];
    c.hu = fx ? -qx : qx;
    c.hv = fy ? -qy : qy;
    return c;
}

__device__ __forceinline__ void put4(double* c, const PC& p)
{
    c[0] = p.he;
    c[1] = p.hu;
    c[2] = p.hv;
    c[3] = p.z;
}

// ---------------------------------------------------------------------------
// Fused face VJPs + gather, one workgroup per tile of kTX x kTY cells.
//
// The workgroup evaluates the VJP of every face whose stencil touches one of
// its cells (x faces of its rows, y faces of its columns) and accumulates
// the results in an LDS image of the tile grown by the stencil reach.  Each
// face adds its stencil slots in separate passes (slot s of every face in
// pass s), so no two threads update one LDS cell in a pass and every cell
// receives its contributions in a fixed order: the sums are deterministic.
// The gather then reads, for each tile cell, its own position and the ghost
// positions filled from it, and applies the ghost-fill and rebuild chain
// rules.  Faces on tile edges are evaluated by both neighbouring tiles
// ((kTX + 3) / kTX and (kTY + 3) / kTY redundancy for the hydrostatic
// stencil).
// ---------------------------------------------------------------------------
constexpr int kTX = 64, kTY = 16;

// Hydrostatic tiles are kHX x kHY cells.  The workgroup first stages every
// cell its faces read -- rows [j0 - 3, j0 + kHY + 2] and columns
// [i0 - 3, i0 + kHX + 2], as the step sees them (rebuilt depth, edge
// terrain, and 1/h, u, v from the momenta with their BC flips) -- once in
// LDS, so each cell is loaded, mapped and divided once instead of by each of
// the 8 faces that read it.  x faces
// accumulate into an image of the tile rows over columns [i0 - 2, i0 + kHX
// + 3), y faces into one of rows [j0 - 2, j0 + kHY + 3) over the tile
// columns (a reflective high side folds ghost n+1 into cell n-2, which may
// be the last cell of the tile before the one holding n-1); x and y faces
// share one flat item list, since they write different images.
constexpr int kHX = 16, kHY = 16, kHThreads = 256;
constexpr int kSX = kHX + 6, kSY = kHY + 6;
constexpr int kAXX = kHX + 5, kAYY = kHY + 5;
struct HShared {
    double cell[5][kSY][kSX];     // he, u, v, z, 1/he (cell_prims)
    double ax[4][kHY][kAXX];
    double ay[4][kAYY][kHX];
};

__device__ __forceinline__ void htile_body(HShared& sh, const State& s, const Geo& g, double dt,
                                           double dx, double dy, const double* __restrict__ lun,
                                           double* __restrict__ lam_new, double* __restrict__ gz,
                                           int bx, int by)
{
    const int64_t N = (int64_t)g.nx * g.ny;
    const int i0 = bx * kHX, j0 = by * kHY;
    const int t = threadIdx.x;
    // stage the cells (indices -2 .. n+1 on each axis: the faces read no
    // further) and clear the images
    for (int k = t; k < kSY * kSX; k += kHThreads) {
        const int ly = k / kSX, lx = k - ly * kSX;
        const int i = i0 - 3 + lx, j = j0 - 3 + ly;
        double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (i >= -2 && i <= g.nx + 1 && j >= -2 && j <= g.ny + 1) {
            const PC p = pcell(s, g, i, j);
            v[0] = p.he;
            v[3] = p.z;
            cell_prims(p.he, p.hu, p.hv, v[4], v[1], v[2]);
        }
        for (int m = 0; m < 5; ++m) sh.cell[m][ly][lx] = v[m];
    }
    for (int k = t; k < 4 * kHY * kAXX; k += kHThreads) (&sh.ax[0][0][0])[k] = 0.0;
    for (int k = t; k < 4 * kAYY * kHX; k += kHThreads) (&sh.ay[0][0][0])[k] = 0.0;
    __syncthreads();
    // x faces f in [i0 - 1, i0 + kHX + 2) of the tile rows, then y faces f in
    // [j0 - 1, j0 + kHY + 2) of the tile columns; the y items start at a
    // multiple of the wavefront size, so each wavefront holds faces of one
    // direction and runs one instance of the VJP
    constexpr int nX = (kHX + 3) * kHY, nY = (kHY + 3) * kHX;
    constexpr int nXw = (nX + 63) / 64 * 64;
    const double rx = dt / dx, ry = dt / dy;
    for (int base = 0; base < nXw + nY; base += kHThreads) {
        const int q = base + t;
        // the face's cells and output adjoints, gathered per direction
        double mL[3] = {0.0, 0.0, 0.0}, mR[3] = {0.0, 0.0, 0.0}, cb[16];
        const double* cs = &sh.cell[0][0][0];   // the face's first stencil cell
        int dir = -1, a0 = 0, a1 = 0;     // image coordinates of slot 0
        if (q < nX) {
            const int row = q / (kHX + 3), fp = q - row * (kHX + 3);
            const int j = j0 + row, f = i0 - 1 + fp;
            if (j < g.ny && f >= 0 && f <= g.nx) {
                cs = &sh.cell[0][row + 3][fp];
                const int64_t cl = (int64_t)j * g.nx + f - 1;
                if (f >= 1)
                    for (int k = 0; k < 3; ++k) mL[k] = -rx * lun[k * N + cl];
                if (f < g.nx)
                    for (int k = 0; k < 3; ++k) mR[k] = rx * lun[k * N + cl + 1];
                dir = 0;
                a0 = row;                // slot k lands on x index f - 2 + k
                a1 = fp - 1;
            }
        } else if (q >= nXw && q < nXw + nY) {
            const int qq = q - nXw;
            const int gp = qq / kHX, cp = qq - gp * kHX;
            const int i = i0 + cp, f = j0 - 1 + gp;
            if (i < g.nx && f >= 0 && f <= g.ny) {
                cs = &sh.cell[0][gp][cp + 3];
                if (f >= 1)
                    for (int k = 0; k < 3; ++k) mL[k] = -ry * lun[k * N + (int64_t)(f - 1) * g.nx + i];
                if (f < g.ny)
                    for (int k = 0; k < 3; ++k) mR[k] = ry * lun[k * N + (int64_t)f * g.nx + i];
                dir = 1;
                a0 = gp - 1;             // slot k lands on y index f - 2 + k
                a1 = cp;
            }
        }
        if (q < nXw) {
            if (dir == 0) hydro_face_vjp<0>(cs, 1, kSY * kSX, mL, mR, cb);
        } else if (dir == 1) {
            hydro_face_vjp<1>(cs, kSX, kSY * kSX, mL, mR, cb);
        }
        for (int k = 0; k < 4; ++k) {
            if (dir == 0) {
                const int lx = a1 + k;
                if (lx >= 0 && lx < kAXX)
                    for (int m = 0; m < 4; ++m) sh.ax[m][a0][lx] += cb[4 * k + m];
            } else if (dir == 1) {
                const int ly = a0 + k;
                if (ly >= 0 && ly < kAYY)
                    for (int m = 0; m < 4; ++m) sh.ay[m][ly][a1] += cb[4 * k + m];
            }
            __syncthreads();
        }
    }
    // gather: the tile cell's own position (x and y images) and the ghost
    // positions filled from it (x ghosts in the x image, y ghosts in the y
    // image), with the ghost-fill and rebuild chain rules; a position's depth
    // was rebuilt from eta where its staged depth is positive
    for (int k = t; k < kHX * kHY; k += kHThreads) {
        const int ci = k % kHX, cj = k / kHX;
        const int i = i0 + ci, j = j0 + cj;
        if (i >= g.nx || j >= g.ny) continue;
        const int64_t c = (int64_t)j * g.nx + i;
        double l0 = lun[c], l1 = lun[N + c], l2 = lun[2 * N + c];   // stored-state path
        double eb = 0.0, zb = 0.0;
        {
            const double he = sh.ax[0][cj][ci + 2] + sh.ay[0][cj + 2][ci];
            l1 += sh.ax[1][cj][ci + 2] + sh.ay[1][cj + 2][ci];
            l2 += sh.ax[2][cj][ci + 2] + sh.ay[2][cj + 2][ci];
            zb += sh.ax[3][cj][ci + 2] + sh.ay[3][cj + 2][ci];
            if (0.0 < sh.cell[0][cj + 3][ci + 3]) { eb += he; zb -= he; }
        }
        if (i <= 1 || i >= g.nx - 2) {
            const int gx[4] = {-2, -1, g.nx, g.nx + 1};
            for (int u = 0; u < 4; ++u) {
                int src, edge;
                bool flip;
                map1(gx[u], g.nx, g.bl0, g.bh0, src, edge, flip);
                if (src != i && edge != i) continue;
                const int lx = gx[u] - (i0 - 2);
                const double he = sh.ax[0][cj][lx];
                const bool wet = 0.0 < sh.cell[0][cj + 3][lx + 1];
                if (src == i) {
                    const double hu = sh.ax[1][cj][lx];
                    l1 += flip ? -hu : hu;
                    l2 += sh.ax[2][cj][lx];
                    if (wet) eb += he;
                }
                if (edge == i) {
                    zb += sh.ax[3][cj][lx];
                    if (wet) zb -= he;
                }
            }
        }
        if (j <= 1 || j >= g.ny - 2) {
            const int gy[4] = {-2, -1, g.ny, g.ny + 1};
            for (int u = 0; u < 4; ++u) {
                int src, edge;
                bool flip;
                map1(gy[u], g.ny, g.bl1, g.bh1, src, edge, flip);
                if (src != j && edge != j) continue;
                const int ly = gy[u] - (j0 - 2);
                const double he = sh.ay[0][ly][ci];
                const bool wet = 0.0 < sh.cell[0][ly + 1][ci + 3];
                if (src == j) {
                    l1 += sh.ay[1][ly][ci];
                    const double hv = sh.ay[2][ly][ci];
                    l2 += flip ? -hv : hv;
                    if (wet) eb += he;
                }
                if (edge == j) {
                    zb += sh.ay[3][ly][ci];
                    if (wet) zb -= he;
                }
            }
        }
        lam_new[c] = l0 + eb;         // eta = h + z
        lam_new[N + c] = l1;
        lam_new[2 * N + c] = l2;
        gz[c] += eb + zb;
    }
}

// Plain Roe: two-cell faces; LDS image covers [i0 - 1, i0 + kTX + 1) (the
// high ghost n folds into n - 1, which is always in the tile holding it).
constexpr int kRPX = kTX + 2, kRPY = kTY + 2;
typedef double RAcc[5][kRPY][kRPX];

__device__ __forceinline__ void rtile_body(RAcc& acc, const State& s, const Geo& g, double dt,
                                           double dx, double dy, const double* __restrict__ lam,
                                           double* __restrict__ lam_new, double* __restrict__ gz,
                                           double* __restrict__ gn, int bx, int by)
{
    const int64_t N = (int64_t)g.nx * g.ny;
    const int i0 = bx * kTX, j0 = by * kTY;
    const int t = threadIdx.x;
    for (int k = t; k < 5 * kRPY * kRPX; k += kThreads) (&acc[0][0][0])[k] = 0.0;
    __syncthreads();
    // un(c) = uo(c) - r (F[xr] - F[xl]) - r (Sp[xr] + Sm[xl])
    for (int dir = 0; dir < 2; ++dir) {
        const double ds = dir == 0 ? dx : dy;
        const double r = dt / ds;
        const int nf = dir == 0 ? (kTX + 1) * kTY : (kTY + 1) * kTX;
        for (int base = 0; base < nf; base += kThreads) {
            const int q = base + t;
            double Lb[5], Rb[5];
            bool have = false;
            int lxL = 0, lyL = 0, lxR = 0, lyR = 0;
            if (q < nf) {
                int i, j, f, n;
                int64_t cl, cr;
                if (dir == 0) {
                    const int row = q / (kTX + 1), fp = q - row * (kTX + 1);
                    j = j0 + row;
                    f = i0 + fp;
                    i = f;
                    n = g.nx;
                    cl = (int64_t)j * g.nx + f - 1;
                    cr = cl + 1;
                    lxL = fp; lyL = row + 1; lxR = fp + 1; lyR = row + 1;
                } else {
                    const int gp = q / kTX, cp = q - gp * kTX;
                    i = i0 + cp;
                    f = j0 + gp;
                    j = f;
                    n = g.ny;
                    cl = (int64_t)(f - 1) * g.nx + i;
                    cr = (int64_t)f * g.nx + i;
                    lxL = cp + 1; lyL = gp; lxR = cp + 1; lyR = gp + 1;
                }
                const bool inside = dir == 0 ? (j < g.ny && f <= g.nx) : (i < g.nx && f <= g.ny);
                if (inside) {
                    const PC a = dir == 0 ? pcell(s, g, f - 1, j) : pcell(s, g, i, f - 1);
                    const PC b = dir == 0 ? pcell(s, g, f, j) : pcell(s, g, i, f);
                    const double L[5] = {a.he, a.hu, a.hv, a.z, a.n};
                    const double R[5] = {b.he, b.hu, b.hv, b.z, b.n};
                    double lL[3] = {0.0, 0.0, 0.0}, lR[3] = {0.0, 0.0, 0.0};
                    if (f >= 1)
                        for (int k = 0; k < 3; ++k) lL[k] = lam[k * N + cl];
                    if (f < n)
                        for (int k = 0; k < 3; ++k) lR[k] = lam[k * N + cr];
                    double Fb[3], Smb[3], Spb[3];
                    for (int k = 0; k < 3; ++k) {
                        Fb[k] = r * lR[k] - r * lL[k];
                        Spb[k] = -r * lL[k];
                        Smb[k] = -r * lR[k];
                    }
                    if (dir == 0) roe_face_vjp<0>(L, R, ds, Fb, Smb, Spb, Lb, Rb);
                    else roe_face_vjp<1>(L, R, ds, Fb, Smb, Spb, Lb, Rb);
                    have = true;
                }
            }
            if (have)
                for (int m = 0; m < 5; ++m) acc[m][lyL][lxL] += Lb[m];
            __syncthreads();
            if (have)
                for (int m = 0; m < 5; ++m) acc[m][lyR][lxR] += Rb[m];
            __syncthreads();
        }
    }
    for (int k = t; k < kTX * kTY; k += kThreads) {
        const int ci = k % kTX, cj = k / kTX;
        const int i = i0 + ci, j = j0 + cj;
        if (i >= g.nx || j >= g.ny) continue;
        const int64_t c = (int64_t)j * g.nx + i;
        // the update starts from the rebuilt depth and the stored momenta
        double l1 = lam[N + c], l2 = lam[2 * N + c];
        double eb = 0.0, zb = 0.0, nb = 0.0;
        {
            const PC p = pcell(s, g, i, j);
            const double he = lam[c] + acc[0][cj + 1][ci + 1];
            l1 += acc[1][cj + 1][ci + 1];
            l2 += acc[2][cj + 1][ci + 1];
            zb += acc[3][cj + 1][ci + 1];
            nb += acc[4][cj + 1][ci + 1];
            if (0.0 < p.x) { eb += he; zb -= he; }
        }
        if (i == 0 || i == g.nx - 1) {
            const int gx[2] = {-1, g.nx};
            for (int u = 0; u < 2; ++u) {
                int src, edge;
                bool flip;
                map1(gx[u], g.nx, g.bl0, g.bh0, src, edge, flip);
                if (src != i && edge != i) continue;
                const int lx = gx[u] - (i0 - 1);
                const PC p = pcell(s, g, gx[u], j);
                const double he = acc[0][cj + 1][lx];
                if (src == i) {
                    const double hu = acc[1][cj + 1][lx];
                    l1 += flip ? -hu : hu;
                    l2 += acc[2][cj + 1][lx];
                    if (0.0 < p.x) eb += he;
                }
                if (edge == i) {
                    zb += acc[3][cj + 1][lx];
                    nb += acc[4][cj + 1][lx];
                    if (0.0 < p.x) zb -= he;
                }
            }
        }
        if (j == 0 || j == g.ny - 1) {
            const int gy[2] = {-1, g.ny};
            for (int u = 0; u < 2; ++u) {
                int src, edge;
                bool flip;
                map1(gy[u], g.ny, g.bl1, g.bh1, src, edge, flip);
                if (src != j && edge != j) continue;
                const int ly = gy[u] - (j0 - 1);
                const PC p = pcell(s, g, i, gy[u]);
                const double he = acc[0][ly][ci + 1];
                if (src == j) {
                    l1 += acc[1][ly][ci + 1];
                    const double hv = acc[2][ly][ci + 1];
                    l2 += flip ? -hv : hv;
                    if (0.0 < p.x) eb += he;
                }
                if (edge == j) {
                    zb += acc[3][ly][ci + 1];
                    nb += acc[4][ly][ci + 1];
                    if (0.0 < p.x) zb -= he;
                }
            }
        }
        lam_new[c] = eb;
        lam_new[N + c] = l1;
        lam_new[2 * N + c] = l2;
        gz[c] += eb + zb;
        gn[c] += nb;
    }
}

// ---------------------------------------------------------------------------
// Batched launches: every kernel below covers all scenarios of a batch in
// one grid.  A workgroup (or thread) finds its scenario by binary search
// over a prefix of work items; scenarios run in lockstep (one step index
// and one segment for all), each with its own buffers, geometry and dt.
// ---------------------------------------------------------------------------
struct Batch {
    const AdjScen* sc;
    const int* start;      // [nscen + 1] prefix of this launch's work items
    int nscen;
    int every, nsteps;
    int t, s0;             // step; first step of its segment (backward)
    int lam_in;            // slot of d.lam holding lambda^{t+1}
};

__device__ __forceinline__ int find_scen(const Batch& b, int item)
{
    int lo = 0, hi = b.nscen - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.start[mid] <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// State t of the segment that starts at s0 (its checkpoint or a recomputed
// state).
__device__ __forceinline__ const double* seg_state(const AdjScen& d, const Batch& b)
{
    const int64_t N3 = 3 * d.N;
    return (b.t == b.s0) ? d.ckpt + (int64_t)(b.s0 / b.every) * N3
                         : d.seg + (int64_t)(b.t - b.s0 - 1) * N3;
}

__device__ __forceinline__ Geo geo_of(const AdjScen& d)
{
    return Geo{d.p.nx, d.p.ny, d.p.bl0, d.p.bl1, d.p.bh0, d.p.bh1};
}

// start: tile prefix
__global__ __launch_bounds__(kHThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void htile_kernel(Batch b)
{
    __shared__ HShared acc;
    const int blk = blockIdx.x;
    const AdjScen& d = b.sc[find_scen(b, blk)];
    const int lt = blk - d.tile0;
    const double* U = seg_state(d, b);
    const State s{U, U + d.N, U + 2 * d.N, d.p.z, d.p.n};
    double* lam_new = d.lam + (int64_t)(1 - b.lam_in) * 3 * d.N;
    htile_body(acc, s, geo_of(d), d.dts[b.t], d.dx, d.dy, d.lun, lam_new, d.gz, lt % d.ntx,
               lt / d.ntx);
}

__global__ __launch_bounds__(kThreads) void rtile_kernel(Batch b)
{
    __shared__ RAcc acc;
    const int blk = blockIdx.x;
    const AdjScen& d = b.sc[find_scen(b, blk)];
    const int lt = blk - d.tile0;
    const double* U = seg_state(d, b);
    const State s{U, U + d.N, U + 2 * d.N, d.p.z, d.p.n};
    const double* lam = d.lam + (int64_t)b.lam_in * 3 * d.N;
    double* lam_new = d.lam + (int64_t)(1 - b.lam_in) * 3 * d.N;
    rtile_body(acc, s, geo_of(d), d.dts[b.t], d.dx, d.dy, lam, lam_new, d.gz, d.gn, lt % d.ntx,
               lt / d.ntx);
}

// Cell-wise kernels: start is a prefix of workgroups, kCellBlocks(N) per
// scenario, each thread striding over its scenario's cells.
constexpr int kCellsPerThread = 8;
inline int cell_blocks(int64_t N)
{
    return (int)((N + (int64_t)kThreads * kCellsPerThread - 1) / ((int64_t)kThreads * kCellsPerThread));
}

#define HX_ADJ_CELLS(d, lb, nb)                                                   \
    const int _blk = blockIdx.x;                                                  \
    const int _s = find_scen(b, _blk);                                            \
    const AdjScen& d = b.sc[_s];                                                  \
    const int lb = _blk - b.start[_s], nb = b.start[_s + 1] - b.start[_s]

// Reverse of Manning friction and the dry clamp of the hydrostatic update
// (HydrostaticRoe.cpp:566-603), cell by cell, from the stored state after
// the flux update: lun = d(output)/d(un)^T lam, gn += d/dn.
__global__ __launch_bounds__(kThreads) void post_kernel(Batch b)
{
    HX_ADJ_CELLS(d, lb, nb);
    const int64_t N = d.N;
    const double* un = d.un + (int64_t)(b.t - b.s0) * 3 * N;
    const double* lam = d.lam + (int64_t)b.lam_in * 3 * N;
    const double dt = d.dts[b.t];
    for (int64_t o = (int64_t)lb * kThreads + threadIdx.x; o < N; o += (int64_t)nb * kThreads) {
        const double u[3] = {un[o], un[N + o], un[2 * N + o]};
        const double ob[3] = {lam[o], lam[N + o], lam[2 * N + o]};
        double ub[3];
        const double dn = hydro_post_vjp(u, d.p.n[o], dt, ob, ub);
        d.lun[o] = ub[0];
        d.lun[N + o] = ub[1];
        d.lun[2 * N + o] = ub[2];
        d.gn[o] += dn;
    }
}

// Start of the backward: lambda = the final-state adjoint (or 0), zero
// gradients.
__global__ __launch_bounds__(kThreads) void init_kernel(Batch b)
{
    HX_ADJ_CELLS(d, lb, nb);
    const int64_t N = d.N;
    double* lam = d.lam + (int64_t)b.lam_in * 3 * N;
    for (int64_t o = (int64_t)lb * kThreads + threadIdx.x; o < N; o += (int64_t)nb * kThreads) {
        for (int k = 0; k < 3; ++k) lam[k * N + o] = d.gfin ? d.gfin[k * N + o] : 0.0;
        d.gz[o] = 0.0;
        d.gn[o] = 0.0;
    }
}

// Forward start: the initial state into checkpoint 0.  Backward end:
// lambda^0 into the caller's state gradient (dir 1).
__global__ __launch_bounds__(kThreads) void copy_kernel(Batch b, int dir)
{
    HX_ADJ_CELLS(d, lb, nb);
    const int64_t N = d.N;
    const double* lam = d.lam + (int64_t)b.lam_in * 3 * N;
    for (int64_t o = (int64_t)lb * kThreads + threadIdx.x; o < N; o += (int64_t)nb * kThreads) {
        if (dir == 0) {
            d.ckpt[o] = d.in_h[o];
            d.ckpt[N + o] = d.in_hu[o];
            d.ckpt[2 * N + o] = d.in_hv[o];
        } else {
            for (int k = 0; k < 3; ++k) d.gstate[k * N + o] = lam[k * N + o];
        }
    }
}

// Depth at the gauges after forward step t; start: gauge prefix, one thread
// per gauge.
__global__ void sample_kernel(Batch b, const int64_t* __restrict__ gauges, int total)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const AdjScen& d = b.sc[find_scen(b, g)];
    const int t = b.t;
    const int64_t N3 = 3 * d.N;
    const double* h;
    if (t == b.nsteps - 1) h = d.out_h;
    else h = ((t + 1) % b.every == 0) ? d.ckpt + (int64_t)((t + 1) / b.every) * N3
                                      : d.seg + (int64_t)((t + 1) & 1) * N3;
    d.series[(int64_t)t * d.ngauge + (g - d.gauge0)] = h[gauges[g]];
}

// Gauge adjoints into lambda_h: one thread per distinct gauge cell of a
// scenario (start: prefix of distinct cells), which sums the series
// adjoints of its gauges in gauge order (deterministic).
__global__ void seed_kernel(Batch b, const int64_t* __restrict__ ucell,
                            const int* __restrict__ uptr, const int* __restrict__ uidx, int total)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= total) return;
    const AdjScen& d = b.sc[find_scen(b, u)];
    if (!d.gs) return;
    double acc = 0.0;
    for (int k = uptr[u]; k < uptr[u + 1]; ++k) acc += d.gs[(int64_t)b.t * d.ngauge + uidx[k]];
    d.lam[(int64_t)b.lam_in * 3 * d.N + ucell[u]] += acc;
}

__global__ void clear_kernel(AdjScen* sc, int nscen)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nscen) return;
    for (int k = 0; k < 3; ++k) slot_clear(&sc[s].slots[k]);
    *sc[s].err = 0;
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace adj
}  // namespace hx

// ===========================================================================
// C-ABI
// ===========================================================================
using namespace hx;
using namespace hx::adj;

struct hx_adjoint {
    hx_ctx* ctx;
    int solver, nscen;
    int nsteps, every, ncp, nseg;
    std::vector<AdjScen> sc;      // host copy of the scenario table
    std::vector<int64_t> N;       // cells per scenario
    std::vector<int> ngauge;
    AdjScen* d_sc;                // device scenario table
    int* d_strip;                 // [nscen + 1] prefixes: strips, tiles,
    int* d_tile;                  //   cell-wise workgroups, gauges,
    int* d_cellblk;               //   distinct gauge cells
    int* d_gauge;
    int* d_ucellp;
    int total_strips, total_tiles, total_cellblk, total_gauges, total_u;
    int64_t* gauges;              // device: every scenario's gauges, scenario by scenario
    int64_t* ucell;               // device: distinct gauge cells, scenario by scenario
    int* uptr;                    // [total_u + 1] into uidx
    int* uidx;                    // local gauge index per entry
    CflSlot* slots;               // 3 per scenario
    int* err;                     // 1 per scenario
    double* dts;                  // [nscen][nsteps]
    double* vel;                  // [nscen][nsteps]
    std::vector<double*> bufs;    // every device allocation of the scenarios
    std::vector<double> hdts;     // host copy of dts
    int ready;                    // a forward completed and the dt schedules are known
};

static void adj_free(hx_adjoint* a)
{
    for (double* b : a->bufs) (void)hipFree(b);
    void* misc[] = {a->d_sc, a->d_strip, a->d_tile, a->d_cellblk, a->d_gauge, a->d_ucellp,
                    a->gauges, a->ucell, a->uptr, a->uidx, a->slots, a->err, a->dts, a->vel};
    for (void* p : misc)
        if (p) (void)hipFree(p);
}

template <typename T>
static bool dev_copy(T** dst, const std::vector<T>& src)
{
    if (src.empty()) return true;
    return hipMalloc((void**)dst, sizeof(T) * src.size()) == hipSuccess &&
           hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice) == hipSuccess;
}

static bool valid_geom(const hx_patch* geom)
{
    if (!geom || geom->nx < 2 || geom->ny < 2 || geom->dtype != HX_F64) return false;
    if (!(geom->dx > 0.0) || !(geom->dy > 0.0)) return false;
    for (int d = 0; d < 2; ++d)
        if ((geom->bc_lo[d] != 0 && geom->bc_lo[d] != 1) ||
            (geom->bc_hi[d] != 0 && geom->bc_hi[d] != 1))
            return false;
    return true;
}

extern "C" int hx_adjoint_batch_create(hx_ctx* c, int solver, int nscen, const hx_patch* geoms,
                                       int nsteps, int ckpt_every, const int* ngauge,
                                       const int64_t* const* gauges, hx_adjoint** out)
{
    if (!c) return HX_E_STATE;
    if (!out) return HX_E_BADARG;
    *out = nullptr;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    if (nscen < 1 || !geoms || nsteps < 1 || ckpt_every < 1 || !ngauge) return HX_E_BADARG;
    for (int s = 0; s < nscen; ++s) {
        if (!valid_geom(&geoms[s]) || ngauge[s] < 0 || (ngauge[s] > 0 && (!gauges || !gauges[s])))
            return HX_E_BADARG;
        const int64_t N = (int64_t)geoms[s].nx * geoms[s].ny;
        if (geoms[s].nx > (1 << 26) || N > ((int64_t)1 << 40)) return HX_E_BADARG;
        for (int g = 0; g < ngauge[s]; ++g)
            if (gauges[s][g] < 0 || gauges[s][g] >= N) return HX_E_BADARG;
    }
    HX_TRY(hipSetDevice(c->device));
    hx_adjoint* a = new (std::nothrow) hx_adjoint();
    if (!a) return HX_E_NOMEM;
    a->ctx = c;
    a->solver = solver;
    a->nscen = nscen;
    a->nsteps = nsteps;
    a->every = std::min(ckpt_every, nsteps);
    a->ncp = (nsteps + a->every - 1) / a->every;
    // the hydrostatic recomputation also runs a segment's last step (for its
    // pre-friction state) and needs a slot for that step's output
    const bool hyd = (solver == HX_SOLVER_HYDROSTATIC_ROE);
    a->nseg = std::max(a->every - (hyd ? 0 : 1), 2);
    a->sc.assign(nscen, AdjScen{});
    a->N.resize(nscen);
    a->ngauge.assign(ngauge, ngauge + nscen);
    std::vector<int> strip(nscen + 1, 0), tile(nscen + 1, 0), cellb(nscen + 1, 0),
        gpre(nscen + 1, 0), upre(nscen + 1, 0);
    std::vector<int64_t> allg, ucell;
    std::vector<int> uptr(1, 0), uidx;
    // one launch marches every scenario's strips: a batch takes one height
    // for their sum, a single scenario its own
    int batch_rows = 0;
    if (nscen > 1) {
        std::vector<int> bx(nscen), by(nscen);
        for (int s = 0; s < nscen; ++s) { bx[s] = geoms[s].nx; by[s] = geoms[s].ny; }
        batch_rows = host::eff_rows_multi(c, solver, HX_F64, bx.data(), by.data(), nscen);
    }
    bool ok = true;
    for (int s = 0; s < nscen && ok; ++s) {
        const hx_patch& gm = geoms[s];
        AdjScen& d = a->sc[s];
        const int64_t N = (int64_t)gm.nx * gm.ny;
        a->N[s] = N;
        d.N = N;
        hx_patch v = gm;
        v.pitch = gm.nx;
        v.h = v.hu = v.hv = v.eta = nullptr;
        v.ng = 0;
        d.p = host::view_of<double>(&v);
        StepArgs<double> sa{};
        sa.p = d.p;
        sa.opitch = gm.nx;
        host::fill_geometry(sa, solver, batch_rows ? batch_rows : host::eff_rows(c, solver, HX_F64, gm.nx, gm.ny),
                            gm.dx, gm.dy);
        d.dx = sa.dx; d.dy = sa.dy; d.rdx = sa.rdx; d.rdy = sa.rdy; d.dx_pow2 = sa.dx_pow2;
        d.rows = sa.rows; d.nsx = sa.nsx; d.nstrips = sa.nstrips;
        d.strip0 = strip[s];
        strip[s + 1] = strip[s] + d.nstrips;
        const int tx = hyd ? kHX : kTX, ty = hyd ? kHY : kTY;
        d.ntx = (gm.nx + tx - 1) / tx;
        d.ntiles = d.ntx * ((gm.ny + ty - 1) / ty);
        d.tile0 = tile[s];
        tile[s + 1] = tile[s] + d.ntiles;
        cellb[s + 1] = cellb[s] + cell_blocks(N);
        d.ngauge = ngauge[s];
        d.gauge0 = gpre[s];
        gpre[s + 1] = gpre[s] + ngauge[s];
        for (int g = 0; g < ngauge[s]; ++g) allg.push_back(gauges[s][g]);
        // group the scenario's gauges by cell (first appearance order)
        std::vector<int64_t> uc;
        std::vector<std::vector<int>> members;
        for (int g = 0; g < ngauge[s]; ++g) {
            size_t u = 0;
            while (u < uc.size() && uc[u] != gauges[s][g]) ++u;
            if (u == uc.size()) {
                uc.push_back(gauges[s][g]);
                members.emplace_back();
            }
            members[u].push_back(g);
        }
        d.u0 = upre[s];
        d.nu = (int)uc.size();
        upre[s + 1] = upre[s] + d.nu;
        for (size_t u = 0; u < uc.size(); ++u) {
            ucell.push_back(uc[u]);
            uidx.insert(uidx.end(), members[u].begin(), members[u].end());
            uptr.push_back((int)uidx.size());
        }
        const size_t S = sizeof(double) * 3 * (size_t)N;
        auto alloc = [&](double** p, size_t bytes) {
            if (!ok) return;
            ok = hipMalloc((void**)p, bytes) == hipSuccess;
            if (ok) a->bufs.push_back(*p);
        };
        alloc(&d.ckpt, S * a->ncp);
        alloc(&d.seg, S * a->nseg);
        alloc(&d.lam, 2 * S);
        if (hyd) {
            alloc(&d.lun, S);
            alloc(&d.un, S * a->every);
        }
    }
    a->total_strips = strip[nscen];
    a->total_tiles = tile[nscen];
    a->total_cellblk = cellb[nscen];
    a->total_gauges = gpre[nscen];
    a->total_u = upre[nscen];
    ok = ok && hipMalloc(&a->slots, sizeof(CflSlot) * 3 * nscen) == hipSuccess &&
         hipMalloc(&a->err, sizeof(int) * nscen) == hipSuccess &&
         hipMalloc(&a->dts, sizeof(double) * (size_t)nscen * nsteps) == hipSuccess &&
         hipMalloc(&a->vel, sizeof(double) * (size_t)nscen * nsteps) == hipSuccess &&
         dev_copy(&a->d_strip, strip) && dev_copy(&a->d_tile, tile) &&
         dev_copy(&a->d_cellblk, cellb) && dev_copy(&a->d_gauge, gpre) &&
         dev_copy(&a->d_ucellp, upre) && dev_copy(&a->gauges, allg) && dev_copy(&a->ucell, ucell) &&
         dev_copy(&a->uptr, uptr) && dev_copy(&a->uidx, uidx) &&
         hipMalloc(&a->d_sc, sizeof(AdjScen) * nscen) == hipSuccess;
    if (ok)
        for (int s = 0; s < nscen; ++s) {
            AdjScen& d = a->sc[s];
            d.slots = a->slots + 3 * s;
            d.err = a->err + s;
            d.dts = a->dts + (size_t)s * nsteps;
            d.vel = a->vel + (size_t)s * nsteps;
        }
    if (!ok) {
        adj_free(a);
        delete a;
        return HX_E_NOMEM;
    }
    a->hdts.assign((size_t)nscen * nsteps, 0.0);
    *out = a;
    return HX_OK;
}

extern "C" int hx_adjoint_create(hx_ctx* c, int solver, const hx_patch* geom, int nsteps,
                                 int ckpt_every, int ngauge, const int64_t* gauges,
                                 hx_adjoint** out)
{
    if (!c) return HX_E_STATE;
    if (!out) return HX_E_BADARG;
    *out = nullptr;
    if (!valid_geom(geom)) return HX_E_BADARG;
    return hx_adjoint_batch_create(c, solver, 1, geom, nsteps, ckpt_every, &ngauge, &gauges, out);
}

extern "C" int hx_adjoint_destroy(hx_adjoint* a)
{
    if (!a) return HX_OK;
    (void)hipSetDevice(a->ctx->device);
    adj_free(a);
    delete a;
    return HX_OK;
}

static Batch batch_of(hx_adjoint* a, const int* start)
{
    Batch b{};
    b.sc = a->d_sc;
    b.start = start;
    b.nscen = a->nscen;
    b.every = a->every;
    b.nsteps = a->nsteps;
    return b;
}

static int upload_table(hx_adjoint* a, hipStream_t hs)
{
    HX_TRY(hipMemcpyAsync(a->d_sc, a->sc.data(), sizeof(AdjScen) * a->nscen, hipMemcpyHostToDevice,
                          hs));
    return HX_OK;
}

extern "C" int hx_adjoint_batch_forward(hx_adjoint* a, const hx_patch* states, double cfl,
                                        double* dt_sched, double* const* series,
                                        void* const* finals, int32_t* steps_done, void* stream)
{
    if (!a) return HX_E_STATE;
    a->ready = 0;
    if (!states || !dt_sched || !finals) return HX_E_BADARG;
    if (steps_done)
        for (int s = 0; s < a->nscen; ++s) steps_done[s] = 0;
    for (int s = 0; s < a->nscen; ++s) {
        const hx_patch* st = &states[s];
        const AdjScen& d = a->sc[s];
        int rc = host::check_patch(st, 0);
        if (rc) return rc;
        if (st->dtype != HX_F64 || st->nx != d.p.nx || st->ny != d.p.ny || st->pitch != d.p.nx ||
            st->eta || st->dx != (double)d.dx || st->dy != (double)d.dy)
            return HX_E_BADARG;
        if (st->bc_lo[0] != d.p.bl0 || st->bc_lo[1] != d.p.bl1 || st->bc_hi[0] != d.p.bh0 ||
            st->bc_hi[1] != d.p.bh1)
            return HX_E_BADARG;
        if (!finals[3 * s] || !finals[3 * s + 1] || !finals[3 * s + 2]) return HX_E_BADARG;
        if (a->ngauge[s] > 0 && (!series || !series[s])) return HX_E_BADARG;
    }
    const bool use_cfl = cfl > 0.0;
    if (!use_cfl)
        for (size_t k = 0; k < (size_t)a->nscen * a->nsteps; ++k)
            if (!(dt_sched[k] >= 0.0)) return HX_E_BADARG;
    hx_ctx* c = a->ctx;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t hs = (hipStream_t)stream;
    for (int s = 0; s < a->nscen; ++s) {
        AdjScen& d = a->sc[s];
        d.p.z = (const double*)states[s].z;
        d.p.n = (const double*)states[s].n;
        d.in_h = (const double*)states[s].h;
        d.in_hu = (const double*)states[s].hu;
        d.in_hv = (const double*)states[s].hv;
        d.out_h = (double*)finals[3 * s];
        d.out_hu = (double*)finals[3 * s + 1];
        d.out_hv = (double*)finals[3 * s + 2];
        d.series = (a->ngauge[s] > 0) ? series[s] : nullptr;
    }
    int rc = upload_table(a, hs);
    if (rc) return rc;
    if (!use_cfl)
        HX_TRY(hipMemcpyAsync(a->dts, dt_sched, sizeof(double) * a->nscen * a->nsteps,
                              hipMemcpyHostToDevice, hs));
    Batch cb = batch_of(a, a->d_cellblk);
    hipLaunchKernelGGL(copy_kernel, dim3(a->total_cellblk), dim3(kThreads), 0, hs, cb, 0);
    AdjBatch fb{};
    fb.sc = a->d_sc;
    fb.strip_start = a->d_strip;
    fb.nscen = a->nscen;
    fb.total = a->total_strips;
    fb.nsteps = a->nsteps;
    fb.every = a->every;
    fb.s0 = -1;
    fb.use_cfl = use_cfl ? 1 : 0;
    fb.cfl = cfl;
    if (use_cfl) {
        // every scenario's slot 0 <- the CFL of its initial state
        hipLaunchKernelGGL(clear_kernel, dim3((a->nscen + 63) / 64), dim3(64), 0, hs, a->d_sc,
                           a->nscen);
        for (int s = 0; s < a->nscen; ++s) {
            hx_patch p0 = states[s];
            rc = host::reduce_cfl_f64(&p0, a->sc[s].slots, hs);
            if (rc) return rc;
        }
    }
    Batch sb = batch_of(a, a->d_gauge);
    for (int k = 0; k < a->nsteps; ++k) {
        fb.t = k;
        fb.ring = k % 3;
        rc = host::launch_adj_batch_f64(c, a->solver, fb, hs);
        if (rc) return rc;
        if (a->total_gauges > 0) {
            sb.t = k;
            hipLaunchKernelGGL(sample_kernel, dim3((a->total_gauges + 63) / 64), dim3(64), 0, hs, sb,
                               (const int64_t*)a->gauges, a->total_gauges);
        }
    }
    HX_TRY(hipGetLastError());
    std::vector<int> err(a->nscen, 0);
    if (use_cfl) {
        HX_TRY(hipMemcpyAsync(err.data(), a->err, sizeof(int) * a->nscen, hipMemcpyDeviceToHost, hs));
        HX_TRY(hipMemcpyAsync(a->hdts.data(), a->dts, sizeof(double) * a->nscen * a->nsteps,
                              hipMemcpyDeviceToHost, hs));
    }
    HX_TRY(hipStreamSynchronize(hs));
    bool collapse = false;
    for (int s = 0; s < a->nscen; ++s) {
        if (steps_done) steps_done[s] = err[s] ? err[s] - 1 : a->nsteps;
        collapse = collapse || err[s];
    }
    if (collapse) return HX_E_DT_COLLAPSE;
    if (use_cfl) std::copy(a->hdts.begin(), a->hdts.end(), dt_sched);
    else std::copy(dt_sched, dt_sched + (size_t)a->nscen * a->nsteps, a->hdts.begin());
    a->ready = 1;
    return HX_OK;
}

extern "C" int hx_adjoint_forward(hx_adjoint* a, const hx_patch* st, double cfl, double* dt_sched,
                                  double* series, void* ho, void* huo, void* hvo,
                                  int32_t* steps_done, void* stream)
{
    if (!a) return HX_E_STATE;
    if (a->nscen != 1) return HX_E_STATE;
    if (!st || !ho || !huo || !hvo) return HX_E_BADARG;
    void* fin[3] = {ho, huo, hvo};
    return hx_adjoint_batch_forward(a, st, cfl, dt_sched, &series, fin, steps_done, stream);
}

extern "C" int hx_adjoint_batch_backward(hx_adjoint* a, const double* const* g_series,
                                         const double* const* g_final, double* const* g_state,
                                         double* const* g_z, double* const* g_n, void* stream)
{
    if (!a) return HX_E_STATE;
    if (!a->ready) return HX_E_STATE;
    if (!g_state || !g_z || !g_n) return HX_E_BADARG;
    for (int s = 0; s < a->nscen; ++s)
        if (!g_state[s] || !g_z[s] || !g_n[s]) return HX_E_BADARG;
    hx_ctx* c = a->ctx;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t hs = (hipStream_t)stream;
    for (int s = 0; s < a->nscen; ++s) {
        AdjScen& d = a->sc[s];
        d.gs = (g_series && a->ngauge[s] > 0) ? g_series[s] : nullptr;
        d.gfin = g_final ? g_final[s] : nullptr;
        d.gstate = g_state[s];
        d.gz = g_z[s];
        d.gn = g_n[s];
    }
    int rc = upload_table(a, hs);
    if (rc) return rc;
    int lam_in = 0;
    Batch cb = batch_of(a, a->d_cellblk);
    cb.lam_in = lam_in;
    hipLaunchKernelGGL(init_kernel, dim3(a->total_cellblk), dim3(kThreads), 0, hs, cb);
    const bool hyd = (a->solver == HX_SOLVER_HYDROSTATIC_ROE);
    AdjBatch fb{};
    fb.sc = a->d_sc;
    fb.strip_start = a->d_strip;
    fb.nscen = a->nscen;
    fb.total = a->total_strips;
    fb.nsteps = a->nsteps;
    fb.every = a->every;
    fb.use_cfl = 0;
    Batch tb = batch_of(a, a->d_tile), ub = batch_of(a, a->d_ucellp);
    for (int cp = a->ncp - 1; cp >= 0; --cp) {
        const int s0 = cp * a->every, s1 = std::min(s0 + a->every, a->nsteps);
        // recompute the segment's states from its checkpoint; the
        // hydrostatic steps also store their states before friction, the
        // segment's last step included (its output lands in a spare slot)
        fb.s0 = s0;
        for (int t = s0; t < (hyd ? s1 : s1 - 1); ++t) {
            fb.t = t;
            rc = host::launch_adj_batch_f64(c, a->solver, fb, hs);
            if (rc) return rc;
        }
        for (int t = s1 - 1; t >= s0; --t) {
            if (g_series && a->total_u > 0) {
                ub.t = t;
                ub.lam_in = lam_in;
                hipLaunchKernelGGL(seed_kernel, dim3((a->total_u + 63) / 64), dim3(64), 0, hs, ub,
                                   (const int64_t*)a->ucell, (const int*)a->uptr,
                                   (const int*)a->uidx, a->total_u);
            }
            if (hyd) {
                cb.t = t;
                cb.s0 = s0;
                cb.lam_in = lam_in;
                hipLaunchKernelGGL(post_kernel, dim3(a->total_cellblk), dim3(kThreads), 0, hs, cb);
            }
            tb.t = t;
            tb.s0 = s0;
            tb.lam_in = lam_in;
            if (hyd) hipLaunchKernelGGL(htile_kernel, dim3(a->total_tiles), dim3(kHThreads), 0, hs, tb);
            else hipLaunchKernelGGL(rtile_kernel, dim3(a->total_tiles), dim3(kThreads), 0, hs, tb);
            lam_in ^= 1;
        }
    }
    cb.lam_in = lam_in;
    hipLaunchKernelGGL(copy_kernel, dim3(a->total_cellblk), dim3(kThreads), 0, hs, cb, 1);
    return host::hip_status(hipGetLastError());
}

extern "C" int hx_adjoint_backward(hx_adjoint* a, const double* g_series, const double* g_final,
                                   double* g_state, double* g_z, double* g_n, void* stream)
{
    if (!a) return HX_E_STATE;
    if (a->nscen != 1) return HX_E_STATE;
    return hx_adjoint_batch_backward(a, &g_series, &g_final, &g_state, &g_z, &g_n, stream);
}
