// hx_io.cpp -- HDF5 input schema of HydroEXA (reader, writer, geometry
// contract) behind include/hydroexa_io.h.  Serial HDF5: the reference opens
// the file with MPI-IO on every rank and reads independent hyperslabs
// (include/io/HDF5_File.H:57-81); on one GPU one process reads.
#include "../../include/hydroexa_io.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

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
