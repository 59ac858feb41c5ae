// binding_check.cpp -- a C++ caller of the C-ABI in include/hydroexa.h, built
// with the host compiler alone (no HIP kernels here) the way a HydroEXA
// solver translation unit would call the library: device memory from the HIP
// runtime API, one hx_patch per box, status codes checked.
//
//   binding_check --version          print the ABI version (no device use)
//   binding_check IN OUT             run the sequence below on IN, write OUT
//
// IN (raw little-endian): int32 nx, ny, solver, nsteps; double dx, dy, cfl,
// dt; then five FP64 planes h, hu, hv, z, n of ny x nx (i fastest).  Sides:
// x low reflective, x high transmissive, y low transmissive, y high
// reflective.
// OUT: the state after hx_advance_level_fluxes with dt (3 planes), its
// (2 ny + 2 nx) x 3 edge fluxes, hx_level_dt's (dt, max|vel|) of that state,
// then the state after a further nsteps of hx_run_level (3 planes) and its
// per-step dt (nsteps doubles).
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hydroexa.h"

namespace {

void die(const char* what, int rc)
{
    std::fprintf(stderr, "%s failed: %d (%s)\n", what, rc, hx_status_string(rc));
    std::exit(1);
}

void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

struct Dev {
    void* p = nullptr;
    explicit Dev(size_t bytes) { hip_ok(hipMalloc(&p, bytes), "hipMalloc"); }
    ~Dev() { hipFree(p); }
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc == 2 && std::strcmp(argv[1], "--version") == 0) {
        std::printf("hydroexa abi %d\n", hx_abi_version());
        return 0;
    }
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s --version | IN OUT\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hdr[4];
    double par[4];
    if (std::fread(hdr, sizeof(hdr), 1, f) != 1 || std::fread(par, sizeof(par), 1, f) != 1) return 2;
    const int nx = hdr[0], ny = hdr[1], solver = hdr[2], nsteps = hdr[3];
    const size_t cells = (size_t)nx * ny, plane = cells * sizeof(double);
    std::vector<double> host(5 * cells);
    if (std::fread(host.data(), sizeof(double), host.size(), f) != host.size()) return 2;
    std::fclose(f);

    hx_ctx* ctx = nullptr;
    int rc = hx_ctx_create(&ctx, 0);
    if (rc != HX_OK) die("hx_ctx_create", rc);

    Dev in(5 * plane), out(3 * plane), alt(3 * plane), ef(3 * (2 * (size_t)ny + 2 * nx) * sizeof(double));
    hip_ok(hipMemcpy(in.p, host.data(), 5 * plane, hipMemcpyHostToDevice), "upload");
    char* base = (char*)in.p;
    hx_patch p{};
    p.nx = nx; p.ny = ny; p.ng = 0; p.dtype = HX_F64; p.pitch = nx;
    p.h = base; p.hu = base + plane; p.hv = base + 2 * plane; p.eta = nullptr;
    p.z = base + 3 * plane; p.n = base + 4 * plane;
    p.dx = par[0]; p.dy = par[1];
    p.bc_lo[0] = HX_BC_REFLECT; p.bc_hi[0] = HX_BC_TRANSMIT;
    p.bc_lo[1] = HX_BC_TRANSMIT; p.bc_hi[1] = HX_BC_REFLECT;

    // one AdvanceLevel with the fine side of the flux register
    char* o = (char*)out.p;
    rc = hx_advance_level_fluxes(ctx, solver, &p, par[3], o, o + plane, o + 2 * plane, nx, ef.p,
                                 nullptr);
    if (rc != HX_OK) die("hx_advance_level_fluxes", rc);
    std::vector<double> res(3 * cells), efh(3 * (2 * (size_t)ny + 2 * nx));
    hip_ok(hipMemcpy(res.data(), out.p, 3 * plane, hipMemcpyDeviceToHost), "download");
    hip_ok(hipMemcpy(efh.data(), ef.p, efh.size() * sizeof(double), hipMemcpyDeviceToHost), "download");

    // ComputeDt of the new state
    hx_patch q = p;
    q.h = o; q.hu = o + plane; q.hv = o + 2 * plane;
    double dt_new = 0.0, vmax = 0.0;
    rc = hx_level_dt(ctx, solver, &q, par[2], &dt_new, &vmax, nullptr);
    if (rc != HX_OK) die("hx_level_dt", rc);

    // nsteps of the fused loop from the new state
    std::vector<double> dth(nsteps > 0 ? nsteps : 1);
    char* a = (char*)alt.p;
    int32_t done = 0;
    rc = hx_run_level(ctx, solver, &q, a, a + plane, a + 2 * plane, par[2], nsteps, dth.data(),
                      nullptr, &done, nullptr);
    if (rc != HX_OK) die("hx_run_level", rc);
    std::vector<double> fin(3 * cells);
    hip_ok(hipMemcpy(fin.data(), (nsteps % 2) ? alt.p : out.p, 3 * plane, hipMemcpyDeviceToHost),
           "download");
    hx_ctx_destroy(ctx);

    FILE* g = std::fopen(argv[2], "wb");
    if (!g) return 2;
    const double dtv[2] = {dt_new, vmax};
    std::fwrite(res.data(), sizeof(double), res.size(), g);
    std::fwrite(efh.data(), sizeof(double), efh.size(), g);
    std::fwrite(dtv, sizeof(double), 2, g);
    std::fwrite(fin.data(), sizeof(double), fin.size(), g);
    std::fwrite(dth.data(), sizeof(double), (size_t)nsteps, g);
    std::fclose(g);
    std::printf("done %d steps\n", (int)done);
    return 0;
}
