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

// Hydrostatic: LDS image covers x in [i0 - 2, i0 + kTX + 3) and likewise in
// y: a reflective high side folds ghost n+1 into cell n-2, which may be the
// last cell of the tile before the one holding n-1.
constexpr int kHPX = kTX + 5, kHPY = kTY + 5;

__global__ __launch_bounds__(kThreads) void htile_kernel(State s, Geo g, double dt, double dx,
                                                         double dy,
                                                         const double* __restrict__ lun,
                                                         double* __restrict__ lam_new,
                                                         double* __restrict__ gz)
{
    __shared__ double acc[4][kHPY][kHPX];
    const int64_t N = (int64_t)g.nx * g.ny;
    const int i0 = blockIdx.x * kTX, j0 = blockIdx.y * kTY;
    const int t = threadIdx.x;
    for (int k = t; k < 4 * kHPY * kHPX; k += kThreads) (&acc[0][0][0])[k] = 0.0;
    __syncthreads();
    // x faces f in [i0 - 1, i0 + kTX + 2) of the tile rows
    {
        const double r = dt / dx;
        constexpr int nf = (kTX + 3) * kTY;
        for (int base = 0; base < nf; base += kThreads) {
            const int q = base + t;
            double cb[16];
            bool have = false;
            int lx0 = 0, ly = 0;
            if (q < nf) {
                const int row = q / (kTX + 3), fp = q - row * (kTX + 3);
                const int j = j0 + row, f = i0 - 1 + fp;
                if (j < g.ny && f >= 0 && f <= g.nx) {
                    double c[16];
                    for (int k = 0; k < 4; ++k) put4(c + 4 * k, pcell(s, g, f - 2 + k, j));
                    double mL[3] = {0.0, 0.0, 0.0}, mR[3] = {0.0, 0.0, 0.0};
                    const int64_t cl = (int64_t)j * g.nx + f - 1;
                    if (f >= 1)
                        for (int k = 0; k < 3; ++k) mL[k] = -r * lun[k * N + cl];
                    if (f < g.nx)
                        for (int k = 0; k < 3; ++k) mR[k] = r * lun[k * N + cl + 1];
                    hydro_face_vjp(c, 1.0, 0.0, mL, mR, cb);
                    have = true;
                    lx0 = fp - 1;        // slot k lands on x index f - 2 + k
                    ly = row + 2;
                }
            }
            for (int k = 0; k < 4; ++k) {
                const int lx = lx0 + k;
                if (have && lx >= 0 && lx < kHPX)
                    for (int m = 0; m < 4; ++m) acc[m][ly][lx] += cb[4 * k + m];
                __syncthreads();
            }
        }
    }
    // y faces f in [j0 - 1, j0 + kTY + 2) of the tile columns
    {
        const double r = dt / dy;
        constexpr int nf = (kTY + 3) * kTX;
        for (int base = 0; base < nf; base += kThreads) {
            const int q = base + t;
            double cb[16];
            bool have = false;
            int ly0 = 0, lx = 0;
            if (q < nf) {
                const int gp = q / kTX, cp = q - gp * kTX;
                const int i = i0 + cp, f = j0 - 1 + gp;
                if (i < g.nx && f >= 0 && f <= g.ny) {
                    double c[16];
                    for (int k = 0; k < 4; ++k) put4(c + 4 * k, pcell(s, g, i, f - 2 + k));
                    double mL[3] = {0.0, 0.0, 0.0}, mR[3] = {0.0, 0.0, 0.0};
                    if (f >= 1)
                        for (int k = 0; k < 3; ++k) mL[k] = -r * lun[k * N + (int64_t)(f - 1) * g.nx + i];
                    if (f < g.ny)
                        for (int k = 0; k < 3; ++k) mR[k] = r * lun[k * N + (int64_t)f * g.nx + i];
                    hydro_face_vjp(c, 0.0, 1.0, mL, mR, cb);
                    have = true;
                    ly0 = gp - 1;
                    lx = cp + 2;
                }
            }
            for (int k = 0; k < 4; ++k) {
                const int ly = ly0 + k;
                if (have && ly >= 0 && ly < kHPY)
                    for (int m = 0; m < 4; ++m) acc[m][ly][lx] += cb[4 * k + m];
                __syncthreads();
            }
        }
    }
    // gather
    for (int k = t; k < kTX * kTY; k += kThreads) {
        const int ci = k % kTX, cj = k / kTX;
        const int i = i0 + ci, j = j0 + cj;
        if (i >= g.nx || j >= g.ny) continue;
        const int64_t c = (int64_t)j * g.nx + i;
        double l0 = lun[c], l1 = lun[N + c], l2 = lun[2 * N + c];   // stored-state path
        double eb = 0.0, zb = 0.0;
        {
            const PC p = pcell(s, g, i, j);
            const double he = acc[0][cj + 2][ci + 2];
            l1 += acc[1][cj + 2][ci + 2];
            l2 += acc[2][cj + 2][ci + 2];
            zb += acc[3][cj + 2][ci + 2];
            if (0.0 < p.x) { eb += he; zb -= he; }
        }
        if (i <= 1 || i >= g.nx - 2) {
            const int gx[4] = {-2, -1, g.nx, g.nx + 1};
            for (int u = 0; u < 4; ++u) {
                int src, edge;
                bool flip;
                map1(gx[u], g.nx, g.bl0, g.bh0, src, edge, flip);
                if (src != i && edge != i) continue;
                const int lx = gx[u] - (i0 - 2);
                const PC p = pcell(s, g, gx[u], j);
                const double he = acc[0][cj + 2][lx];
                if (src == i) {
                    const double hu = acc[1][cj + 2][lx];
                    l1 += flip ? -hu : hu;
                    l2 += acc[2][cj + 2][lx];
                    if (0.0 < p.x) eb += he;
                }
                if (edge == i) {
                    zb += acc[3][cj + 2][lx];
                    if (0.0 < p.x) zb -= he;
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
                const PC p = pcell(s, g, i, gy[u]);
                const double he = acc[0][ly][ci + 2];
                if (src == j) {
                    l1 += acc[1][ly][ci + 2];
                    const double hv = acc[2][ly][ci + 2];
                    l2 += flip ? -hv : hv;
                    if (0.0 < p.x) eb += he;
                }
                if (edge == j) {
                    zb += acc[3][ly][ci + 2];
                    if (0.0 < p.x) zb -= he;
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

__global__ __launch_bounds__(kThreads) void rtile_kernel(State s, Geo g, double dt, double dx,
                                                         double dy,
                                                         const double* __restrict__ lam,
                                                         double* __restrict__ lam_new,
                                                         double* __restrict__ gz,
                                                         double* __restrict__ gn)
{
    __shared__ double acc[5][kRPY][kRPX];
    const int64_t N = (int64_t)g.nx * g.ny;
    const int i0 = blockIdx.x * kTX, j0 = blockIdx.y * kTY;
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
                    roe_face_vjp(L, R, ds, dir == 0 ? 1.0 : 0.0, dir == 0 ? 0.0 : 1.0, Fb, Smb,
                                 Spb, Lb, Rb);
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
// Small kernels
// ---------------------------------------------------------------------------
// Reverse of Manning friction and the dry clamp of the hydrostatic update
// (HydrostaticRoe.cpp:566-603), cell by cell, from the stored state after
// the flux update: lun = d(output)/d(un)^T lam, gn += d/dn.
__global__ __launch_bounds__(kThreads) void post_kernel(const double* __restrict__ un,
                                                        const double* __restrict__ n,
                                                        const double* __restrict__ lam,
                                                        double* __restrict__ lun,
                                                        double* __restrict__ gn, double dt,
                                                        int64_t N)
{
    for (int64_t o = blockIdx.x * (int64_t)kThreads + threadIdx.x; o < N;
         o += (int64_t)gridDim.x * kThreads) {
        const double u[3] = {un[o], un[N + o], un[2 * N + o]};
        const double ob[3] = {lam[o], lam[N + o], lam[2 * N + o]};
        double ub[3];
        const double dn = hydro_post_vjp(u, n[o], dt, ob, ub);
        lun[o] = ub[0];
        lun[N + o] = ub[1];
        lun[2 * N + o] = ub[2];
        gn[o] += dn;
    }
}

__global__ void fill_kernel(double* p, int64_t n, double v)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * blockDim.x)
        p[k] = v;
}

__global__ void sample_kernel(const double* h, const int64_t* gauges, double* out, int ng)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ng) out[g] = h[gauges[g]];
}

// Gauge adjoints into lambda_h: one thread per distinct gauge cell, which
// sums the series adjoints of its gauges in gauge order (deterministic).
__global__ void seed_kernel(double* lam_h, const int64_t* ucell, const int* uptr,
                            const int* uidx, const double* gs, int nu)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    double acc = 0.0;
    for (int k = uptr[u]; k < uptr[u + 1]; ++k) acc += gs[uidx[k]];
    lam_h[ucell[u]] += acc;
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
    int solver, nx, ny;
    double dx, dy;
    int bc_lo[2], bc_hi[2];
    int nsteps, every, ncp, nseg, ngauge;
    int64_t N;
    int64_t* gauges;   // device, ngauge flat cell indices
    int64_t* ucell;    // device, distinct gauge cells
    int* uptr;         // device, nu + 1: gauges of distinct cell u are
    int* uidx;         //   uidx[uptr[u] .. uptr[u+1]) (device, ngauge)
    int nu;
    double* ckpt;      // ncp states (3N each): the state at steps 0, every, 2 every, ...
    double* seg;       // nseg states: recomputed states inside a segment
    double* un;        // hydrostatic: every states after the flux update and
                       // before friction, stored by the segment's recomputation
    double* lam;       // 2 states: adjoint ping-pong
    double* lun;       // 1 state: adjoint before friction (hydrostatic)
    double* hist;      // 2 nsteps: dt and max|vel| of the forward's CFL steps
    std::vector<double> dts;
    const double* z;
    const double* n;
    int ready;         // a forward completed and the dt schedule is known
};

static void adj_free(hx_adjoint* a)
{
    double* bufs[] = {a->ckpt, a->seg, a->lam, a->lun, a->un, a->hist};
    for (double* b : bufs)
        if (b) (void)hipFree(b);
    if (a->gauges) (void)hipFree(a->gauges);
    if (a->ucell) (void)hipFree(a->ucell);
    if (a->uptr) (void)hipFree(a->uptr);
    if (a->uidx) (void)hipFree(a->uidx);
}

extern "C" int hx_adjoint_create(hx_ctx* c, int solver, const hx_patch* geom, int nsteps,
                                 int ckpt_every, int ngauge, const int64_t* gauges,
                                 hx_adjoint** out)
{
    if (!c) return HX_E_STATE;
    if (!out) return HX_E_BADARG;
    *out = nullptr;
    if (solver != HX_SOLVER_ROE && solver != HX_SOLVER_HYDROSTATIC_ROE) return HX_E_BADARG;
    if (!geom || geom->nx < 2 || geom->ny < 2 || geom->dtype != HX_F64) return HX_E_BADARG;
    if (!(geom->dx > 0.0) || !(geom->dy > 0.0)) return HX_E_BADARG;
    for (int d = 0; d < 2; ++d)
        if ((geom->bc_lo[d] != 0 && geom->bc_lo[d] != 1) ||
            (geom->bc_hi[d] != 0 && geom->bc_hi[d] != 1))
            return HX_E_BADARG;
    if (nsteps < 1 || ckpt_every < 1 || ngauge < 0 || (ngauge > 0 && !gauges)) return HX_E_BADARG;
    const int64_t N = (int64_t)geom->nx * geom->ny;
    for (int g = 0; g < ngauge; ++g)
        if (gauges[g] < 0 || gauges[g] >= N) return HX_E_BADARG;
    HX_TRY(hipSetDevice(c->device));
    hx_adjoint* a = new (std::nothrow) hx_adjoint();
    if (!a) return HX_E_NOMEM;
    a->ctx = c;
    a->solver = solver;
    a->nx = geom->nx;
    a->ny = geom->ny;
    a->dx = geom->dx;
    a->dy = geom->dy;
    for (int d = 0; d < 2; ++d) {
        a->bc_lo[d] = geom->bc_lo[d];
        a->bc_hi[d] = geom->bc_hi[d];
    }
    a->nsteps = nsteps;
    a->every = std::min(ckpt_every, nsteps);
    a->ncp = (nsteps + a->every - 1) / a->every;
    // the hydrostatic recomputation also runs a segment's last step (for its
    // pre-friction state) and needs a slot for that step's output
    a->nseg = std::max(a->every - (solver == HX_SOLVER_HYDROSTATIC_ROE ? 0 : 1), 2);
    a->ngauge = ngauge;
    a->N = N;
    const size_t S = sizeof(double) * 3 * (size_t)N;
    bool ok = hipMalloc(&a->ckpt, S * a->ncp) == hipSuccess &&
              hipMalloc(&a->seg, S * a->nseg) == hipSuccess &&
              hipMalloc(&a->lam, 2 * S) == hipSuccess &&
              hipMalloc(&a->hist, sizeof(double) * 2 * (size_t)a->nsteps) == hipSuccess;
    if (ok && solver == HX_SOLVER_HYDROSTATIC_ROE)
        ok = hipMalloc(&a->lun, S) == hipSuccess && hipMalloc(&a->un, S * a->every) == hipSuccess;
    if (ok && ngauge > 0) {
        // group gauges by cell (first appearance order) for the seed kernel
        std::vector<int64_t> ucell;
        std::vector<std::vector<int>> members;
        for (int gi = 0; gi < ngauge; ++gi) {
            size_t u = 0;
            while (u < ucell.size() && ucell[u] != gauges[gi]) ++u;
            if (u == ucell.size()) {
                ucell.push_back(gauges[gi]);
                members.emplace_back();
            }
            members[u].push_back(gi);
        }
        std::vector<int> uptr(1, 0), uidx;
        for (auto& m : members) {
            uidx.insert(uidx.end(), m.begin(), m.end());
            uptr.push_back((int)uidx.size());
        }
        a->nu = (int)ucell.size();
        ok = hipMalloc(&a->gauges, sizeof(int64_t) * ngauge) == hipSuccess &&
             hipMemcpy(a->gauges, gauges, sizeof(int64_t) * ngauge, hipMemcpyHostToDevice) ==
                 hipSuccess &&
             hipMalloc(&a->ucell, sizeof(int64_t) * a->nu) == hipSuccess &&
             hipMemcpy(a->ucell, ucell.data(), sizeof(int64_t) * a->nu, hipMemcpyHostToDevice) ==
                 hipSuccess &&
             hipMalloc(&a->uptr, sizeof(int) * (a->nu + 1)) == hipSuccess &&
             hipMemcpy(a->uptr, uptr.data(), sizeof(int) * (a->nu + 1), hipMemcpyHostToDevice) ==
                 hipSuccess &&
             hipMalloc(&a->uidx, sizeof(int) * ngauge) == hipSuccess &&
             hipMemcpy(a->uidx, uidx.data(), sizeof(int) * ngauge, hipMemcpyHostToDevice) ==
                 hipSuccess;
    }
    if (!ok) {
        adj_free(a);
        delete a;
        return HX_E_NOMEM;
    }
    a->dts.assign(nsteps, 0.0);
    *out = a;
    return HX_OK;
}

extern "C" int hx_adjoint_destroy(hx_adjoint* a)
{
    if (!a) return HX_OK;
    (void)hipSetDevice(a->ctx->device);
    adj_free(a);
    delete a;
    return HX_OK;
}

static double* state_ptr(hx_adjoint* a, double* base, int k) { return base + (size_t)k * 3 * a->N; }

static void set_io(StepArgs<double>& s, const double* in, int64_t N, double* ho, double* huo,
                   double* hvo)
{
    s.p.h = in;
    s.p.hu = in + N;
    s.p.hv = in + 2 * N;
    s.ho = ho;
    s.huo = huo;
    s.hvo = hvo;
}

extern "C" int hx_adjoint_forward(hx_adjoint* a, const hx_patch* st, double cfl, double* dt_sched,
                                  double* series, void* ho, void* huo, void* hvo,
                                  int32_t* steps_done, void* stream)
{
    if (!a) return HX_E_STATE;
    a->ready = 0;
    if (steps_done) *steps_done = 0;
    int rc = host::check_patch(st, 0);
    if (rc) return rc;
    if (st->dtype != HX_F64 || st->nx != a->nx || st->ny != a->ny || st->pitch != a->nx ||
        st->eta || st->dx != a->dx || st->dy != a->dy)
        return HX_E_BADARG;
    for (int d = 0; d < 2; ++d)
        if (st->bc_lo[d] != a->bc_lo[d] || st->bc_hi[d] != a->bc_hi[d]) return HX_E_BADARG;
    if (!dt_sched || !ho || !huo || !hvo || (a->ngauge > 0 && !series)) return HX_E_BADARG;
    const bool use_cfl = cfl > 0.0;
    if (!use_cfl)
        for (int s = 0; s < a->nsteps; ++s)
            if (!(dt_sched[s] >= 0.0)) return HX_E_BADARG;
    hx_ctx* c = a->ctx;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t hs = (hipStream_t)stream;
    const int64_t N = a->N;
    a->z = (const double*)st->z;
    a->n = (const double*)st->n;
    double* c0 = a->ckpt;
    HX_TRY(hipMemcpyAsync(c0, st->h, sizeof(double) * N, hipMemcpyDeviceToDevice, hs));
    HX_TRY(hipMemcpyAsync(c0 + N, st->hu, sizeof(double) * N, hipMemcpyDeviceToDevice, hs));
    HX_TRY(hipMemcpyAsync(c0 + 2 * N, st->hv, sizeof(double) * N, hipMemcpyDeviceToDevice, hs));
    StepArgs<double> s{};
    s.p = host::view_of<double>(st);
    s.opitch = a->nx;
    host::fill_geometry(s, a->solver, host::eff_rows(c, a->solver, HX_F64, a->nx, a->ny), a->dx, a->dy);
    c->mark.valid = 0;       // the context's slots now follow this rollout
    if (use_cfl) {
        rc = host::clear_slots(c->slots, 3, c->err, hs);
        if (rc) return rc;
        hx_patch p0 = *st;
        p0.h = c0;
        p0.hu = c0 + N;
        p0.hv = c0 + 2 * N;
        rc = host::reduce_cfl_f64(&p0, c->slots, hs);
        if (rc) return rc;
        s.cfl = cfl;
        s.err = c->err;
        s.dt_hist = a->hist;
        s.vel_hist = a->hist + a->nsteps;
    }
    const double* in = c0;
    for (int k = 0; k < a->nsteps; ++k) {
        double *oh, *ohu, *ohv;
        if (k == a->nsteps - 1) {
            oh = (double*)ho;
            ohu = (double*)huo;
            ohv = (double*)hvo;
        } else {
            double* o = ((k + 1) % a->every == 0) ? state_ptr(a, a->ckpt, (k + 1) / a->every)
                                                  : state_ptr(a, a->seg, (k + 1) & 1);
            oh = o;
            ohu = o + N;
            ohv = o + 2 * N;
        }
        set_io(s, in, N, oh, ohu, ohv);
        if (use_cfl) {
            s.rd = c->slots + (k % 3);
            s.acc = c->slots + ((k + 1) % 3);
            s.clr = c->slots + ((k + 2) % 3);
            s.step = k;
        } else {
            s.rd = nullptr;
            s.dt_host = dt_sched[k];
        }
        rc = host::launch_step_f64(c, a->solver, s, hs);
        if (rc) return rc;
        if (a->ngauge > 0) {
            hipLaunchKernelGGL(sample_kernel, dim3((a->ngauge + 63) / 64), dim3(64), 0, hs, oh,
                               a->gauges, series + (size_t)k * a->ngauge, a->ngauge);
            HX_TRY(hipGetLastError());
        }
        in = oh;   // oh, ohu, ohv are planes of one state except for the last step
    }
    if (use_cfl) {
        int err = 0;
        HX_TRY(hipMemcpyAsync(&err, c->err, sizeof(int), hipMemcpyDeviceToHost, hs));
        HX_TRY(hipMemcpyAsync(a->dts.data(), a->hist, sizeof(double) * a->nsteps,
                              hipMemcpyDeviceToHost, hs));
        HX_TRY(hipStreamSynchronize(hs));
        if (err) {
            if (steps_done) *steps_done = err - 1;
            return HX_E_DT_COLLAPSE;
        }
        std::copy(a->dts.begin(), a->dts.end(), dt_sched);
    } else {
        std::copy(dt_sched, dt_sched + a->nsteps, a->dts.begin());
    }
    if (steps_done) *steps_done = a->nsteps;
    a->ready = 1;
    return HX_OK;
}

// One adjoint step: lam (adjoint of the step's output) -> lam_new (adjoint
// of its input U), accumulating the bed and roughness gradients.  un: the
// step's state after the flux update (hydrostatic), stored by the forward
// recomputation.
static int adj_step(hx_adjoint* a, const double* U, const double* un, double dt,
                    const double* lam, double* lam_new, double* gz, double* gn, hipStream_t hs)
{
    const int64_t N = a->N;
    State s{U, U + N, U + 2 * N, a->z, a->n};
    Geo g{a->nx, a->ny, a->bc_lo[0], a->bc_lo[1], a->bc_hi[0], a->bc_hi[1]};
    const dim3 tiles((unsigned)((a->nx + kTX - 1) / kTX), (unsigned)((a->ny + kTY - 1) / kTY));
    if (a->solver == HX_SOLVER_HYDROSTATIC_ROE) {
        const unsigned pb = (unsigned)std::min<int64_t>(nblocks(N), 8192);
        hipLaunchKernelGGL(post_kernel, dim3(pb), dim3(kThreads), 0, hs, un, (const double*)a->n,
                           lam, a->lun, gn, dt, N);
        hipLaunchKernelGGL(htile_kernel, tiles, dim3(kThreads), 0, hs, s, g, dt, a->dx, a->dy,
                           (const double*)a->lun, lam_new, gz);
    } else {
        hipLaunchKernelGGL(rtile_kernel, tiles, dim3(kThreads), 0, hs, s, g, dt, a->dx, a->dy,
                           lam, lam_new, gz, gn);
    }
    return host::hip_status(hipGetLastError());
}

extern "C" int hx_adjoint_backward(hx_adjoint* a, const double* g_series, const double* g_final,
                                   double* g_state, double* g_z, double* g_n, void* stream)
{
    if (!a) return HX_E_STATE;
    if (!a->ready) return HX_E_STATE;
    if (!g_state || !g_z || !g_n) return HX_E_BADARG;
    hx_ctx* c = a->ctx;
    HX_TRY(hipSetDevice(c->device));
    hipStream_t hs = (hipStream_t)stream;
    const int64_t N = a->N;
    double* lam = a->lam;
    double* lam2 = a->lam + 3 * N;
    const unsigned fb = (unsigned)std::min<int64_t>(nblocks(3 * N), 8192);
    if (g_final)
        HX_TRY(hipMemcpyAsync(lam, g_final, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, hs));
    else
        hipLaunchKernelGGL(fill_kernel, dim3(fb), dim3(kThreads), 0, hs, lam, 3 * N, 0.0);
    hipLaunchKernelGGL(fill_kernel, dim3(fb), dim3(kThreads), 0, hs, g_z, N, 0.0);
    hipLaunchKernelGGL(fill_kernel, dim3(fb), dim3(kThreads), 0, hs, g_n, N, 0.0);
    HX_TRY(hipGetLastError());
    StepArgs<double> s{};
    hx_patch geom{};
    geom.nx = a->nx;
    geom.ny = a->ny;
    geom.dtype = HX_F64;
    geom.pitch = a->nx;
    geom.z = a->z;
    geom.n = a->n;
    geom.dx = a->dx;
    geom.dy = a->dy;
    for (int d = 0; d < 2; ++d) {
        geom.bc_lo[d] = a->bc_lo[d];
        geom.bc_hi[d] = a->bc_hi[d];
    }
    s.p = host::view_of<double>(&geom);
    s.opitch = a->nx;
    host::fill_geometry(s, a->solver, host::eff_rows(c, a->solver, HX_F64, a->nx, a->ny), a->dx, a->dy);
    for (int cp = a->ncp - 1; cp >= 0; --cp) {
        const int s0 = cp * a->every, s1 = std::min(s0 + a->every, a->nsteps);
        auto state = [&](int t) -> double* {
            return (t == s0) ? state_ptr(a, a->ckpt, cp) : state_ptr(a, a->seg, t - s0 - 1);
        };
        // recompute the segment's states from its checkpoint; the
        // hydrostatic steps also store their states before friction, the
        // segment's last step included (its output lands in a spare slot)
        const bool hyd = (a->solver == HX_SOLVER_HYDROSTATIC_ROE);
        for (int t = s0; t < (hyd ? s1 : s1 - 1); ++t) {
            double* o = state(t + 1);
            set_io(s, state(t), N, o, o + N, o + 2 * N);
            s.dt_host = a->dts[t];
            s.uno = hyd ? a->un + (size_t)(t - s0) * 3 * N : nullptr;
            s.uno_plane = N;
            int rc = host::launch_step_f64(c, a->solver, s, hs);
            s.uno = nullptr;
            if (rc) return rc;
        }
        for (int t = s1 - 1; t >= s0; --t) {
            if (g_series && a->ngauge > 0) {
                hipLaunchKernelGGL(seed_kernel, dim3((a->nu + 63) / 64), dim3(64), 0, hs, lam,
                                   a->ucell, a->uptr, a->uidx,
                                   g_series + (size_t)t * a->ngauge, a->nu);
            }
            const double* un = a->un ? a->un + (size_t)(t - s0) * 3 * N : nullptr;
            int rc = adj_step(a, state(t), un, a->dts[t], lam, lam2, g_z, g_n, hs);
            if (rc) return rc;
            std::swap(lam, lam2);
        }
    }
    HX_TRY(hipMemcpyAsync(g_state, lam, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, hs));
    return HX_OK;
}
