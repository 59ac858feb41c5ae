// hx_io.cpp -- HydroEXA's file formats behind include/hydroexa_io.h: the HDF5
// input schema (reader, writer, geometry contract) and the native AMReX
// plotfile writer.  Serial HDF5: the reference opens the file with MPI-IO on
// every rank and reads independent hyperslabs (include/io/HDF5_File.H:57-81);
// on one GPU one process reads.
#include "../../include/hydroexa_io.h"

#include <hdf5.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>

namespace {

// Closes an HDF5 handle on scope exit.
struct H {
    hid_t id;
    herr_t (*close)(hid_t);
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H()
    {
        if (id >= 0) close(id);
    }
    H(const H&) = delete;
    H& operator=(const H&) = delete;
};

void quiet()
{
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
}

double attr_or(hid_t dset, const char* name, double fallback, bool* found = nullptr)
{
    if (found) *found = false;
    if (H5Aexists(dset, name) <= 0) return fallback;
    H a(H5Aopen(dset, name, H5P_DEFAULT), H5Aclose);
    if (a.id < 0) return fallback;
    double v = fallback;
    if (H5Aread(a.id, H5T_NATIVE_DOUBLE, &v) < 0) return fallback;
    if (found) *found = true;
    return v;
}

int put_attr(hid_t dset, const char* name, double v)
{
    H sp(H5Screate(H5S_SCALAR), H5Sclose);
    H a(H5Acreate2(dset, name, H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.id < 0) return HXIO_E_WRITE;
    return H5Awrite(a.id, H5T_NATIVE_DOUBLE, &v) < 0 ? HXIO_E_WRITE : HXIO_OK;
}

int put_dataset(hid_t file, const char* name, const double* data, int nc, int ny, int nx,
                double dx, double dy, double x_ll, double y_ll)
{
    const hsize_t dims[3] = {(hsize_t)nc, (hsize_t)ny, (hsize_t)nx};
    H sp(H5Screate_simple(3, dims, nullptr), H5Sclose);
    H d(H5Dcreate2(file, name, H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
        H5Dclose);
    if (d.id < 0) return HXIO_E_WRITE;
    if (H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        return HXIO_E_WRITE;
    int rc = put_attr(d.id, "dx", dx);
    if (!rc) rc = put_attr(d.id, "dy", dy);
    if (!rc) rc = put_attr(d.id, "x_ll", x_ll);
    if (!rc) rc = put_attr(d.id, "y_ll", y_ll);
    return rc;
}

void set_msg(char* msg, int msglen, const char* text)
{
    if (msg && msglen > 0) {
        std::strncpy(msg, text, (size_t)msglen - 1);
        msg[msglen - 1] = '\0';
    }
}

}  // namespace

extern "C" int hxio_read_metadata(const char* path, const char* dataset, hxio_meta* out)
{
    if (!path || !dataset || !out) return HXIO_E_BADARG;
    quiet();
    H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) return HXIO_E_OPEN;
    H d(H5Dopen2(f.id, dataset, H5P_DEFAULT), H5Dclose);
    if (d.id < 0) return HXIO_E_DATASET;
    H sp(H5Dget_space(d.id), H5Sclose);
    const int nd = H5Sget_simple_extent_ndims(sp.id);
    if (nd != 2 && nd != 3) return HXIO_E_SHAPE;
    hsize_t dims[3] = {0, 0, 0};
    H5Sget_simple_extent_dims(sp.id, dims, nullptr);
    hxio_meta m{};
    m.ndims = nd;
    // the trailing two dimensions are spatial (include/io/HDF5_File.H:116-130)
    if (nd == 3) {
        m.ncomp = (int32_t)dims[0];
        m.ny = (int32_t)dims[1];
        m.nx = (int32_t)dims[2];
    } else {
        m.ncomp = 1;
        m.ny = (int32_t)dims[0];
        m.nx = (int32_t)dims[1];
    }
    bool has_dy = false;
    m.dx = attr_or(d.id, "dx", 1.0);
    m.dy = attr_or(d.id, "dy", m.dx, &has_dy);   // dy defaults to dx (IOHandler.cpp:257-266)
    m.x_ll = attr_or(d.id, "x_ll", 0.0);
    m.y_ll = attr_or(d.id, "y_ll", 0.0);
    m.x_hi = m.x_ll + m.nx * m.dx;
    m.y_hi = m.y_ll + m.ny * m.dy;
    *out = m;
    return HXIO_OK;
}

extern "C" int hxio_read_components(const char* path, const char* dataset, int first_comp,
                                    int ncomp, int i_lo, int j_lo, int nx, int ny, double nodata,
                                    double* out)
{
    if (!path || !dataset || !out || ncomp < 1 || nx < 1 || ny < 1 || first_comp < 0)
        return HXIO_E_BADARG;
    quiet();
    H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) return HXIO_E_OPEN;
    H d(H5Dopen2(f.id, dataset, H5P_DEFAULT), H5Dclose);
    if (d.id < 0) return HXIO_E_DATASET;
    H fs(H5Dget_space(d.id), H5Sclose);
    if (H5Sget_simple_extent_ndims(fs.id) != 3) return HXIO_E_SHAPE;   // IOHandler.cpp:143
    hsize_t dims[3];
    H5Sget_simple_extent_dims(fs.id, dims, nullptr);
    if ((hsize_t)(first_comp + ncomp) > dims[0]) return HXIO_E_SHAPE;
    const int fny = (int)dims[1], fnx = (int)dims[2];
    const size_t plane = (size_t)nx * ny;
    // cells outside the file box read nodata (IOHandler.cpp:206-218)
    std::fill(out, out + plane * ncomp, nodata);
    const int vi0 = std::max(i_lo, 0), vj0 = std::max(j_lo, 0);
    const int vi1 = std::min(i_lo + nx, fnx), vj1 = std::min(j_lo + ny, fny);
    if (vi0 >= vi1 || vj0 >= vj1) return HXIO_OK;
    const hsize_t foff[3] = {(hsize_t)first_comp, (hsize_t)vj0, (hsize_t)vi0};
    const hsize_t cnt[3] = {(hsize_t)ncomp, (hsize_t)(vj1 - vj0), (hsize_t)(vi1 - vi0)};
    if (H5Sselect_hyperslab(fs.id, H5S_SELECT_SET, foff, nullptr, cnt, nullptr) < 0)
        return HXIO_E_READ;
    const hsize_t mdims[3] = {(hsize_t)ncomp, (hsize_t)ny, (hsize_t)nx};
    H ms(H5Screate_simple(3, mdims, nullptr), H5Sclose);
    const hsize_t moff[3] = {0, (hsize_t)(vj0 - j_lo), (hsize_t)(vi0 - i_lo)};
    if (H5Sselect_hyperslab(ms.id, H5S_SELECT_SET, moff, nullptr, cnt, nullptr) < 0)
        return HXIO_E_READ;
    if (H5Dread(d.id, H5T_NATIVE_DOUBLE, ms.id, fs.id, H5P_DEFAULT, out) < 0) return HXIO_E_READ;
    return HXIO_OK;
}

extern "C" int hxio_write_input(const char* path, const double* terrain, const double* fluid,
                                int nx, int ny, double dx, double dy, double x_ll, double y_ll)
{
    if (!path || !terrain || !fluid || nx < 1 || ny < 1) return HXIO_E_BADARG;
    // square cells, as PyHydroEXA's writer requires (hdf5.py:107-111)
    if (std::fabs(dx - dy) > 1e-10 * std::max(std::fabs(dx), std::fabs(dy))) return HXIO_E_BADARG;
    quiet();
    H f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) return HXIO_E_OPEN;
    int rc = put_dataset(f.id, "terrain", terrain, 2, ny, nx, dx, dy, x_ll, y_ll);
    if (!rc) rc = put_dataset(f.id, "fluid", fluid, 3, ny, nx, dx, dy, x_ll, y_ll);
    return rc;
}

extern "C" int hxio_check_geometry(const hxio_meta* m, int ref, int bf, hxio_geometry* out,
                                   char* msg, int msglen)
{
    if (!m || !out) return HXIO_E_BADARG;
    set_msg(msg, msglen, "");
    auto fail = [&](const char* text) {
        set_msg(msg, msglen, text);
        return HXIO_E_GEOMETRY;
    };
    // src/HydroEXA.cpp:115-202, in the reference's order
    if (!(bf > 0)) return fail("amr.blocking_factor must be greater than zero.");
    if (!(ref >= 0)) return fail("amr.terrain_ref_lev must be non-negative.");
    if (!(m->nx > 0 && m->ny > 0)) return fail("HDF5 terrain dimensions must be positive.");
    if (!(m->dx > 0.0 && m->dy > 0.0)) return fail("HDF5 terrain cell spacing must be positive.");
    if (ref > 30) return fail("amr.terrain_ref_lev must be non-negative.");
    const int scale = 1 << ref;
    const double gtol = 1.e-12 * std::max(1.0, std::max(std::fabs(m->dx), std::fabs(m->dy)));
    if (!(std::fabs(m->dx - m->dy) <= gtol))
        return fail("HydroEXA requires square HDF5 cells: dx must equal dy.");
    const double cell = 0.5 * (m->dx + m->dy);
    if (m->nx % scale != 0) return fail("HDF5 global_nx must be divisible by 2^amr.terrain_ref_lev.");
    if (m->ny % scale != 0) return fail("HDF5 global_ny must be divisible by 2^amr.terrain_ref_lev.");
    const int bnx = m->nx / scale, bny = m->ny / scale;
    if (!(bnx > 0 && bny > 0)) return fail("Derived level-0 grid dimensions must be positive.");
    if (bnx % bf != 0)
        return fail("Level-0 nx is incompatible with amr.blocking_factor. HydroEXA will not pad "
                    "the physical domain.");
    if (bny % bf != 0)
        return fail("Level-0 ny is incompatible with amr.blocking_factor. HydroEXA will not pad "
                    "the physical domain.");
    const double l0 = cell * (double)scale;
    const double Lx = m->x_hi - m->x_ll, Ly = m->y_hi - m->y_ll;
    if (!(Lx > 0.0 && Ly > 0.0)) return fail("HDF5 physical domain must have positive extents.");
    const double etol = 1.e-10 * std::max(1.0, std::max(std::fabs(Lx), std::fabs(Ly)));
    if (!(std::fabs((double)bnx * l0 - Lx) <= etol))
        return fail("HDF5 x extent is inconsistent with global_nx, terrain_ref_lev, and dx.");
    if (!(std::fabs((double)bny * l0 - Ly) <= etol))
        return fail("HDF5 y extent is inconsistent with global_ny, terrain_ref_lev, and dy.");
    hxio_geometry g{};
    g.base_nx = bnx;
    g.base_ny = bny;
    g.refinement_scale = scale;
    g.file_cell_size = cell;
    g.level0_dx = l0;
    g.level0_dy = l0;
    g.prob_lo[0] = m->x_ll;
    g.prob_lo[1] = m->y_ll;
    g.prob_hi[0] = m->x_hi;
    g.prob_hi[1] = m->y_hi;
    *out = g;
    return HXIO_OK;
}

// ---------------------------------------------------------------------------
// Native AMReX plotfile (the layout amrex::WriteMultiLevelPlotfile writes:
// plotfile Header, then per level Level_<l>/Cell_H and Level_<l>/Cell_D_00000).
// ---------------------------------------------------------------------------
namespace {

std::string box_str(int ilo, int jlo, int ihi, int jhi)
{
    return "((" + std::to_string(ilo) + "," + std::to_string(jlo) + ") (" + std::to_string(ihi) +
           "," + std::to_string(jhi) + ") (0,0))";
}

// A double as an ostream with the given precision prints it (%g style).
std::string gfmt(double v, int prec)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.*g", prec, v);
    return buf;
}

// VisMF writes min/max in scientific notation with precision 15.
std::string efmt(double v)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.15e", v);
    return buf;
}

bool make_dir(const std::string& d)
{
    size_t pos = 0;
    do {
        pos = d.find('/', pos + 1);
        const std::string part = d.substr(0, pos);
        if (!part.empty()) {
            struct stat st;
            if (::stat(part.c_str(), &st) != 0) {
                if (::mkdir(part.c_str(), 0755) != 0 && errno != EEXIST) return false;
            } else if (!S_ISDIR(st.st_mode)) {
                return false;
            }
        }
    } while (pos != std::string::npos);
    return true;
}

}  // namespace

extern "C" int hxio_write_plotfile(const char* dir, int nlevels, const hxio_plot_level* levels,
                                   int ncomp, const char* const* varnames, double time,
                                   const double* prob_lo, const double* prob_hi,
                                   const int32_t* ref_ratio)
{
    if (!dir || nlevels < 1 || !levels || ncomp < 1 || !varnames || !prob_lo || !prob_hi ||
        (nlevels > 1 && !ref_ratio))
        return HXIO_E_BADARG;
    for (int l = 0; l < nlevels; ++l) {
        if (levels[l].nboxes < 1 || !levels[l].boxes || !levels[l].data) return HXIO_E_BADARG;
        for (int b = 0; b < levels[l].nboxes; ++b) {
            const int32_t* bx = levels[l].boxes + 4 * b;
            if (bx[2] < bx[0] || bx[3] < bx[1] || !levels[l].data[b]) return HXIO_E_BADARG;
        }
    }
    const std::string root(dir);
    if (!make_dir(root)) return HXIO_E_OPEN;
    const int finest = nlevels - 1;
    std::string h;
    h += "HyperCLaw-V1.1\n";
    h += std::to_string(ncomp) + "\n";
    for (int c = 0; c < ncomp; ++c) h += std::string(varnames[c]) + "\n";
    h += "2\n";
    h += gfmt(time, 17) + "\n";
    h += std::to_string(finest) + "\n";
    h += gfmt(prob_lo[0], 17) + " " + gfmt(prob_lo[1], 17) + " \n";
    h += gfmt(prob_hi[0], 17) + " " + gfmt(prob_hi[1], 17) + " \n";
    for (int l = 0; l < finest; ++l) h += std::to_string(ref_ratio[l]) + " ";
    h += "\n";
    for (int l = 0; l < nlevels; ++l) {
        const int32_t* d = levels[l].domain;
        h += box_str(d[0], d[1], d[2], d[3]) + " ";
    }
    h += "\n";
    for (int l = 0; l < nlevels; ++l) h += std::to_string(levels[l].step) + " ";
    h += "\n";
    for (int l = 0; l < nlevels; ++l)
        h += gfmt(levels[l].dx, 17) + " " + gfmt(levels[l].dy, 17) + " \n";
    h += "0\n";   // Cartesian coordinates
    h += "0\n";   // boundary width
    for (int l = 0; l < nlevels; ++l) {
        const hxio_plot_level& L = levels[l];
        h += std::to_string(l) + " " + std::to_string(L.nboxes) + " " + gfmt(time, 17) + "\n";
        h += std::to_string(L.step) + "\n";
        for (int b = 0; b < L.nboxes; ++b) {
            const int32_t* bx = L.boxes + 4 * b;
            const double xlo = prob_lo[0] + (bx[0] - L.domain[0]) * L.dx;
            const double xhi = prob_lo[0] + (bx[2] + 1 - L.domain[0]) * L.dx;
            const double ylo = prob_lo[1] + (bx[1] - L.domain[1]) * L.dy;
            const double yhi = prob_lo[1] + (bx[3] + 1 - L.domain[1]) * L.dy;
            h += gfmt(xlo, 17) + " " + gfmt(xhi, 17) + "\n";
            h += gfmt(ylo, 17) + " " + gfmt(yhi, 17) + "\n";
        }
        h += "Level_" + std::to_string(l) + "/Cell\n";

        // data file: one FAB per box, each a text header then the doubles
        const std::string ldir = root + "/Level_" + std::to_string(l);
        if (!make_dir(ldir)) return HXIO_E_OPEN;
        FILE* fd = std::fopen((ldir + "/Cell_D_00000").c_str(), "wb");
        if (!fd) return HXIO_E_OPEN;
        std::vector<long> offs(L.nboxes);
        std::vector<double> mn((size_t)L.nboxes * ncomp), mx((size_t)L.nboxes * ncomp);
        for (int b = 0; b < L.nboxes; ++b) {
            const int32_t* bx = L.boxes + 4 * b;
            const size_t n = (size_t)(bx[2] - bx[0] + 1) * (bx[3] - bx[1] + 1);
            offs[b] = std::ftell(fd);
            const std::string fh = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" +
                                   box_str(bx[0], bx[1], bx[2], bx[3]) + " " +
                                   std::to_string(ncomp) + "\n";
            if (std::fwrite(fh.data(), 1, fh.size(), fd) != fh.size() ||
                std::fwrite(L.data[b], sizeof(double), n * ncomp, fd) != n * ncomp) {
                std::fclose(fd);
                return HXIO_E_WRITE;
            }
            for (int c = 0; c < ncomp; ++c) {
                const double* p = L.data[b] + c * n;
                double lo = p[0], hi = p[0];
                for (size_t k = 1; k < n; ++k) {
                    lo = std::min(lo, p[k]);
                    hi = std::max(hi, p[k]);
                }
                mn[(size_t)b * ncomp + c] = lo;
                mx[(size_t)b * ncomp + c] = hi;
            }
        }
        if (std::fclose(fd) != 0) return HXIO_E_WRITE;
        // VisMF header (version 1, NFiles)
        std::string m;
        m += "1\n1\n" + std::to_string(ncomp) + "\n0\n";
        m += "(" + std::to_string(L.nboxes) + " 0\n";
        for (int b = 0; b < L.nboxes; ++b) {
            const int32_t* bx = L.boxes + 4 * b;
            m += box_str(bx[0], bx[1], bx[2], bx[3]) + "\n";
        }
        m += ")\n";
        m += std::to_string(L.nboxes) + "\n";
        for (int b = 0; b < L.nboxes; ++b)
            m += "FabOnDisk: Cell_D_00000 " + std::to_string(offs[b]) + "\n";
        m += "\n";
        for (int pass = 0; pass < 2; ++pass) {
            const std::vector<double>& v = pass == 0 ? mn : mx;
            m += std::to_string(L.nboxes) + "," + std::to_string(ncomp) + "\n";
            for (int b = 0; b < L.nboxes; ++b) {
                for (int c = 0; c < ncomp; ++c) m += efmt(v[(size_t)b * ncomp + c]) + ",";
                m += "\n";
            }
            m += "\n";
        }
        FILE* fm = std::fopen((ldir + "/Cell_H").c_str(), "w");
        if (!fm) return HXIO_E_OPEN;
        const bool okm = std::fwrite(m.data(), 1, m.size(), fm) == m.size();
        if (std::fclose(fm) != 0 || !okm) return HXIO_E_WRITE;
    }
    FILE* fh = std::fopen((root + "/Header").c_str(), "w");
    if (!fh) return HXIO_E_OPEN;
    const bool okh = std::fwrite(h.data(), 1, h.size(), fh) == h.size();
    if (std::fclose(fh) != 0 || !okh) return HXIO_E_WRITE;
    return HXIO_OK;
}
