/*
 * hydroexa_io.h -- C-ABI of libhydroexa_io.so: the HDF5 input schema of
 * HydroEXA (read, write) and its geometry contract.  Host code only; the
 * library links the serial HDF5 C library.
 *
 * Schema (PyHydroEXA/src/pyhydroexa/preprocessing/hdf5.py:1-135):
 *   dataset "terrain": 3-D (2, ny, nx) -- z, Manning roughness
 *   dataset "fluid":   3-D (3, ny, nx) -- h, hu, hv
 *   attributes per dataset: "dx", "dy", "x_ll", "y_ll" (doubles)
 *   NODATA sentinel -9999.0
 * Arrays are component-major, row-major in (y, x): element (c, j, i) at
 * c*ny*nx + j*nx + i.
 */
#ifndef HYDROEXA_IO_H
#define HYDROEXA_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HXIO_OK          0
#define HXIO_E_BADARG    2
#define HXIO_E_OPEN     10   /* file cannot be opened or created          */
#define HXIO_E_DATASET  11   /* dataset missing                           */
#define HXIO_E_SHAPE    12   /* dataset rank or component count mismatch  */
#define HXIO_E_READ     13
#define HXIO_E_WRITE    14
#define HXIO_E_GEOMETRY 15   /* geometry contract violated (see message)  */

typedef struct hxio_meta {
    int32_t nx, ny;          /* spatial extents (last two dimensions)      */
    int32_t ndims;           /* rank as stored, 2 or 3                     */
    int32_t ncomp;           /* leading extent of a 3-D dataset, else 1    */
    double dx, dy;           /* attribute "dx"; "dy" defaults to dx        */
    double x_ll, y_ll;       /* lower-left corner                          */
    double x_hi, y_hi;       /* x_ll + nx dx, y_ll + ny dy                 */
} hxio_meta;

/*
 * Dataset extents and spatial attributes.  Replaces
 * IOHandler::ReadHDF5Metadata (src/io/IOHandler.cpp:230-307) and
 * HDF5Reader::Dataset (include/io/HDF5_File.H:99-135), same defaults:
 * dx = 1, dy = dx when absent, x_ll = y_ll = 0.
 */
int hxio_read_metadata(const char *path, const char *dataset, hxio_meta *out);

/*
 * Components [first_comp, first_comp + ncomp) of a 3-D dataset over the
 * box i in [i_lo, i_lo + nx), j in [j_lo, j_lo + ny) (file index space;
 * the box may extend past the file), into out (ncomp x ny x nx doubles).
 * Cells outside the file receive nodata.  Replaces
 * IOHandler::ReadHDF5HyperslabComponents (src/io/IOHandler.cpp:127-227).
 */
int hxio_read_components(const char *path, const char *dataset, int first_comp, int ncomp,
                         int i_lo, int j_lo, int nx, int ny, double nodata, double *out);

/*
 * Write an input file in the schema above.  Replaces PyHydroEXA's
 * write_hdf5 (PyHydroEXA/src/pyhydroexa/preprocessing/hdf5.py:51-135),
 * including its checks: matching extents and square cells
 * (|dx - dy| <= 1e-10 max(|dx|, |dy|)).
 */
int hxio_write_input(const char *path, const double *terrain, const double *fluid, int nx,
                     int ny, double dx, double dy, double x_ll, double y_ll);

typedef struct hxio_geometry {
    int32_t base_nx, base_ny;   /* level-0 cells                           */
    int32_t refinement_scale;   /* 2^terrain_ref_lev                       */
    int32_t pad;
    double file_cell_size;      /* 0.5 (dx + dy)                           */
    double level0_dx, level0_dy;
    double prob_lo[2], prob_hi[2];
} hxio_geometry;

/*
 * The geometry contract of HydroEXA::Initialize (src/HydroEXA.cpp:95-205):
 * square cells, file extents divisible by 2^terrain_ref_lev, level-0
 * extents divisible by the blocking factor, and a physical extent equal to
 * the cell count times the level-0 cell size.  On violation returns
 * HXIO_E_GEOMETRY and writes the reference's assertion text to msg.
 */
int hxio_check_geometry(const hxio_meta *meta, int terrain_ref_lev, int blocking_factor,
                        hxio_geometry *out, char *msg, int msglen);

/*
 * One level of a plotfile: nboxes boxes {ilo, jlo, ihi, jhi} in the level's
 * index space, and per box ncomp planes of its cells (component-major, i
 * fastest), host memory.
 */
typedef struct hxio_plot_level {
    int32_t nboxes;
    int32_t domain[4];            /* {0, 0, nx-1, ny-1}                     */
    int32_t step;                 /* level step count                        */
    const int32_t *boxes;         /* nboxes x 4                              */
    const double *const *data;    /* nboxes pointers                         */
    double dx, dy;
} hxio_plot_level;

/*
 * Native AMReX plotfile (HyperCLaw-V1.1 Header, VisMF Cell_H version 1,
 * one Cell_D_00000 data file per level), as amrex::WriteMultiLevelPlotfile
 * writes it.  Replaces WriteNativePlotfile (src/io/NativePlotFile.cpp:19-97),
 * which calls it with varnames h_fluid, hu_momentum, hv_momentum,
 * z_bathymetry, extra_comp_4 and the directory name prefix + the iteration
 * in 5 digits.  ref_ratio: nlevels-1 ratios.  Creates dir.
 */
int hxio_write_plotfile(const char *dir, int nlevels, const hxio_plot_level *levels, int ncomp,
                        const char *const *varnames, double time, const double *prob_lo,
                        const double *prob_hi, const int32_t *ref_ratio);

#ifdef __cplusplus
}
#endif

#endif /* HYDROEXA_IO_H */
