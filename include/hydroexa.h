/*
 * hydroexa.h -- C-ABI of the MI355X-native per-level shallow-water advance.
 *
 * These entry points replace the per-level solver calls of HydroEXA
 * (jkhansell/HydroEXA); each declaration names the reference interface it
 * stands in for.  Plain pointers and sizes only: device pointers are HIP
 * device memory, host pointers are marked as such, and a stream argument is
 * a hipStream_t passed as void* (NULL = the default stream).
 *
 * Status codes replace the reference's process-fatal errors
 * (std::runtime_error, amrex::Error, amrex::Abort); no exception crosses
 * this boundary.
 */
#ifndef HYDROEXA_H
#define HYDROEXA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HX_ABI_VERSION 1

/* Status codes. */
#define HX_OK             0
#define HX_E_BADARG       2  /* invalid descriptor, size or code            */
#define HX_E_DT_COLLAPSE  3  /* dt estimate < 1e-12 (Roe.cpp:88-92,
                                HydrostaticRoe.cpp:77-80)                   */
#define HX_E_HIP          4  /* a HIP runtime call failed                   */
#define HX_E_NOMEM        5  /* device allocation failed                    */
#define HX_E_STATE        6  /* context used out of order                   */

/* Solver kinds, as selected in src/state/AmrMeshState_Init.cpp:175-184. */
#define HX_SOLVER_ROE             0
#define HX_SOLVER_HYDROSTATIC_ROE 1

/* Boundary codes per side.  0 and 1 are the input-file codes of
 * include/boundaries/BCFill.H:49-82 (0 reflective, 1 transmissive); the
 * library applies them while loading the halo.  HX_BC_GHOST means the
 * caller has filled that side's ghost cells in memory (an AMReX FillPatch,
 * a neighbouring box), and the kernels read them as stored. */
#define HX_BC_REFLECT   0
#define HX_BC_TRANSMIT  1
#define HX_BC_GHOST    (-1)

/* Element types. */
#define HX_F64 0
#define HX_F32 1

/* Arithmetic builds of the fused advance.  EXACT: correctly rounded
 * divisions, no FMA contraction; FP64 results equal the reference's
 * operation order bit for bit.  FAST: shared Newton-refined reciprocals and
 * FMA contraction, each quotient within ~2 ulp. */
#define HX_ARITH_EXACT 0
#define HX_ARITH_FAST  1

/*
 * One box of one level, structure of arrays.  Mirrors the per-FAB view the
 * reference builds in MFIter loops (Array4 of U with comps h, hu, hv and of
 * the terrain with comps z, n; src/solvers/Roe.cpp:327-352).
 *
 * Every field pointer addresses valid cell (0,0); element (i,j) is at
 * ptr[j*pitch + i], and i,j may reach -ng .. n-1+ng on sides coded
 * HX_BC_GHOST.  All fields share pitch and type.
 */
typedef struct hx_patch {
    int32_t nx, ny;         /* valid cells                                  */
    int32_t ng;             /* ghost layers stored on HX_BC_GHOST sides     */
    int32_t dtype;          /* HX_F64 or HX_F32                             */
    int64_t pitch;          /* elements between rows                        */
    const void *h, *hu, *hv;/* state at t^n                                 */
    const void *eta;        /* optional free surface; NULL: eta = h + z     */
    const void *z, *n;      /* bed elevation and Manning roughness          */
    double dx, dy;          /* cell size                                    */
    int32_t bc_lo[2];       /* per direction: HX_BC_* at the low side       */
    int32_t bc_hi[2];       /* per direction: HX_BC_* at the high side      */
} hx_patch;

typedef struct hx_ctx hx_ctx;

/* Library context: owns the small device scratch the kernels reduce into
 * (CFL slots, dt history).  One context per host thread; calls that share
 * a context must be ordered on one stream.  (No reference counterpart: the
 * reference allocates its temporaries per call, Roe.cpp:291-303.) */
int hx_ctx_create(hx_ctx **out, int device);
int hx_ctx_destroy(hx_ctx *ctx);
int hx_abi_version(void);
const char *hx_status_string(int status);

/*
 * One AdvanceLevel of one box: U_new = U_old - dt * div(fluxes).
 * Replaces SolverVariant::compute_fluxes (include/solvers/SolverVariant.H:81-90)
 * -> Roe::compute_fluxes_Impl (src/solvers/Roe.cpp:250-434) or
 * HydrostaticRoe::compute_fluxes_Impl (src/solvers/HydrostaticRoe.cpp:400-606),
 * fused into one kernel.  Output arrays (device, out_pitch) receive the new
 * h, hu, hv of the valid cells; they must not alias the inputs.  The
 * reference swaps U_new/U_old inside the call; here the caller swaps.
 * Reflux face fluxes are not produced (hx_advance_level_fluxes does).
 */
int hx_advance_level(hx_ctx *ctx, int solver, const hx_patch *in, double dt,
                     void *h_out, void *hu_out, void *hv_out, int64_t out_pitch,
                     void *stream);

/*
 * hx_advance_level, plus the effective fluxes of the box's boundary faces as
 * the box's own cells took them in the same kernel: the fine side of a flux
 * register (src/solvers/Roe.cpp:544-566, the fr_as_fine branch of the
 * face-oriented selector: D- = F - S- on the low faces, whose cell lies on
 * the high side, D+ = F + S+ on the high faces; hydrostatic 0.5 FL + 0.5 FR,
 * src/solvers/HydrostaticRoe.cpp:612-627).  edge_flux (device, patch dtype,
 * not aliasing anything else): (2 ny + 2 nx) rows of 3 values (h, hu, hv
 * components, x-y frame), unscaled: rows [0, ny) the x-low faces (i = 0)
 * from j = 0 up, [ny, 2 ny) the x-high faces (i = nx), [2 ny, 2 ny + nx) the
 * y-low faces (j = 0) from i = 0 right, [2 ny + nx, 2 ny + 2 nx) the y-high
 * faces (j = ny).  The reference hands these to FluxRegister::FineAdd with
 * scale dt dA; they equal what hx_face_fluxes re-evaluates for the same
 * faces and sides.
 */
int hx_advance_level_fluxes(hx_ctx *ctx, int solver, const hx_patch *in, double dt,
                            void *h_out, void *hu_out, void *hv_out, int64_t out_pitch,
                            void *edge_flux, void *stream);

/*
 * CFL timestep of one level.  Replaces compute_dt_Impl
 * (src/solvers/Roe.cpp:8-98, src/solvers/HydrostaticRoe.cpp:7-86):
 * *dt_out (host) = cfl * min over cells, *max_vel_out (host) = max |u|,|v|.
 * Synchronises the stream.  Returns HX_E_DT_COLLAPSE (and still fills the
 * outputs) when the estimate before cfl is below 1e-12.
 */
int hx_level_dt(hx_ctx *ctx, int solver, const hx_patch *in, double cfl,
                double *dt_out, double *max_vel_out, void *stream);

/*
 * Gradient tagging.  Replaces SolverVariant::tag_cells
 * (include/solvers/SolverVariant.H:139-156) -> Roe::tag_cells_Impl
 * (src/solvers/Roe.cpp:100-181) or HydrostaticRoe::tag_cells_Impl
 * (src/solvers/HydrostaticRoe.cpp:88-171).  thresh (host) = {h_grad,
 * umag_grad, z_grad} of the level (PhysicsParameters, include/utils/Params.H:103-111).
 * tags (device, nx*ny bytes, row pitch nx) receive 1 where the reference
 * sets TagBox::SET and 0 elsewhere.  Ghost semantics: physical sides read
 * the halo each component's BCRec would produce; HX_BC_GHOST sides read
 * memory.
 */
int hx_tag_level(hx_ctx *ctx, int solver, const hx_patch *in, double time,
                 const double *thresh, uint8_t *tags, void *stream);

/*
 * NODATA-aware restriction.  Replaces amrex::masked_average_down
 * (include/utils/MaskedAverageDown.H:124-127, src/utils/MaskedAverageDown.cpp:12-63).
 * fine: ncomp planes of nyf rows x nxf (device, dense); crse: ncomp planes
 * of (nyf/ry) x (nxf/rx); nvalid (device, optional): valid-child count per
 * coarse cell and component.  dxf, dyf: fine cell size (the volume weight).
 */
int hx_masked_average_down(hx_ctx *ctx, int dtype, int nxf, int nyf, int ncomp,
                           const void *fine, int rx, int ry, double dxf, double dyf,
                           double nodata, void *crse, int32_t *nvalid, void *stream);

/*
 * Base-grid pruning predicate.  Replaces AmrMeshState::PostProcessBaseGrids
 * (src/state/AmrMeshState_AMR.cpp:318-405).  boxes (host): nbox x
 * {ilo, jlo, ihi, jhi} in level-0 cells; z (device): static bed of nys x nxs
 * cells at 2^scale_log2 times level-0 resolution.  keep/has_nodata (host,
 * nbox bytes) and *nkeep (host) as in the reference, including its
 * keep-everything fallback when no box holds a valid cell.  Synchronises.
 */
int hx_prune_boxes(hx_ctx *ctx, int dtype, int nbox, const int32_t *boxes, int nxs,
                   int nys, const void *z, int scale_log2, uint8_t *keep,
                   uint8_t *has_nodata, int32_t *nkeep, void *stream);

/*
 * Conservation diagnostics of one level in one fused reduction.  Replaces
 * the per-level reduction of AmrMeshState::ComputeMassDiagnostics
 * (src/state/AmrMeshState_Utils.cpp:213-343): over valid cells whose
 * roughness is not NODATA (-9999) and h > 0, out (host, 5 doubles) =
 * {sum h dA, sum hu dA, sum hv dA, sum 0.5 (hu^2 + hv^2)/h dA, min h}
 * (min h = DBL_MAX when no cell qualifies).  Sums are compensated and
 * deterministic; totals over levels and ranks are the caller's, as in the
 * reference.  Synchronises the stream.
 */
int hx_conservation(hx_ctx *ctx, const hx_patch *level, double *out, void *stream);

/*
 * Measurement utility with no reference counterpart: the streaming ceiling of
 * the fused advance's access pattern.  One elementwise pass over n FP64
 * elements that reads in[0..4] and writes out[0..2]
 * (out0 = in0 + s, out1 = in1 * s, out2 = in2 - s, s = in3 + in4), enqueued
 * on stream.  Device pointers; returns HX_E_BADARG for n <= 0 or a null
 * pointer.
 */
int hx_stream_probe(hx_ctx *ctx, int64_t n, const void *const *in, void *const *out, void *stream);

/*
 * Step-driven single-level time loop with the CFL reduction fused into the
 * advance: nsteps x (ComputeDt; AdvanceLevel).  Replaces the loop body of
 * HydroEXA::Compute (src/HydroEXA.cpp:260-271) for one level with a step
 * count instead of max_time (src/state/AmrMeshState_Physics.cpp:13-19,104-108).
 * The dt of step k is reduced on the device by step k-1 (no host sync per
 * step).  state: patch whose h/hu/hv hold t^0 on entry (all sides physical);
 * alt_h/alt_hu/alt_hv: a second state buffer with the same pitch.  On
 * return the final state is in state's buffers if nsteps is even and in the
 * alt buffers if odd.  dt_hist / vel_hist (host, optional, nsteps each)
 * receive dt and max|vel| per step.  Returns HX_E_DT_COLLAPSE if a step's
 * estimate fell below 1e-12 (*steps_done, host, says how many completed).
 * Allocates nothing: the context's history holds 4096 steps and a longer
 * run synchronises once per 4096 steps.
 */
int hx_run_level(hx_ctx *ctx, int solver, const hx_patch *state, void *alt_h,
                 void *alt_hu, void *alt_hv, double cfl, int nsteps, double *dt_hist,
                 double *vel_hist, int32_t *steps_done, void *stream);

/*
 * hx_run_level for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc);
 * it has no counterpart in the reference.  Same arguments, kernels and
 * results as hx_run_level, but the CFL slots, error flag and dt history live
 * in the library's static per-device scratch (about 152 KiB, resolved when
 * the first context on the device is created) instead of the context's
 * buffers, and nothing is allocated or freed.  At most 8192 steps per call
 * (HX_E_BADARG beyond).  Every call on one GPU shares that scratch, which
 * keeps the last call's contents: calls that overlap on one GPU (several
 * host threads or streams) race.  Does not start or end a rollout that
 * hx_run_level_resume could continue.
 */
int hx_run_level_noalloc(hx_ctx *ctx, int solver, const hx_patch *state, void *alt_h,
                         void *alt_hu, void *alt_hv, double cfl, int nsteps, double *dt_hist,
                         double *vel_hist, int32_t *steps_done, void *stream);

/*
 * Continue the rollout that the last successful hx_run_level or
 * hx_run_level_resume on this context finished: the first step's dt comes
 * from the CFL estimate of the final state that its last fused advance
 * reduced (the reference's ComputeDt at the top of the next loop iteration,
 * src/HydroEXA.cpp:260-271), so a run split into calls (output between
 * them) costs exactly its fused advances and gives the bits of one long
 * call.  state must name that call's final buffers (the alt buffers after
 * an odd step count) with the same solver, geometry, terrain and BC codes,
 * unchanged since; otherwise HX_E_STATE.  hx_adjoint_forward on the same
 * context ends the rollout (it reuses the slots).  Other arguments and
 * results as hx_run_level.  No reference counterpart beyond the loop.
 */
int hx_run_level_resume(hx_ctx *ctx, int solver, const hx_patch *state, void *alt_h,
                        void *alt_hu, void *alt_hv, double cfl, int nsteps, double *dt_hist,
                        double *vel_hist, int32_t *steps_done, void *stream);

/*
 * Launch-shape control for the fused advance (the autotuner's knobs): rows
 * marched per wavefront and the minimum waves per SIMD the register
 * allocator must leave room for (1, 3 or 4); 0 selects the measured default
 * when the advance is launched (3 waves per SIMD for FP64; a strip height
 * chosen from the level's size, the solver and the kernel's occupancy).
 * Results do not depend on either.  No reference counterpart (AMReX picks
 * launch shapes itself).
 */
int hx_set_tuning(hx_ctx *ctx, int solver, int dtype, int rows, int occupancy);

/*
 * Select the arithmetic build of the fused advance (HX_ARITH_EXACT, the
 * default, or HX_ARITH_FAST).  No reference counterpart; the reference's
 * arithmetic is what HX_ARITH_EXACT reproduces.
 */
int hx_set_arith(hx_ctx *ctx, int mode);

/*
 * Batched ensemble: many independent levels (flood scenarios of uneven
 * sizes) advanced together, one launch per step for the whole batch, each
 * scenario with its own CFL-limited dt (the reference's per-level
 * compute_dt + AdvanceLevel, src/state/AmrMeshState_Physics.cpp:13-19,153-172,
 * applied per scenario).  No batching exists in the reference; each scenario
 * matches hx_run_level on it alone bit for bit.
 *
 * scen: nscen patches (all sides physical, one dtype, no eta), h/hu/hv
 * holding t^0; alt: 3*nscen device pointers, buffer B (h, hu, hv) of each
 * scenario, same pitch.  After hx_ensemble_run the state is in the patches'
 * buffers if nsteps is even, in alt otherwise.  The ensemble keeps the
 * pointers; the caller keeps the memory alive.
 */
typedef struct hx_ensemble hx_ensemble;
int hx_ensemble_create(hx_ctx *ctx, int solver, int dtype, int nscen, const hx_patch *scen,
                       void *const *alt, hx_ensemble **out);
int hx_ensemble_destroy(hx_ensemble *ens);

/* rows marched per wavefront, waves per SIMD (0 = defaults) and the
 * schedule: 0 = one wavefront per strip, 1 = persistent wavefronts pulling
 * strips from an atomic work queue. */
int hx_ensemble_set_tuning(hx_ensemble *ens, int rows, int occupancy, int persistent);

/*
 * nsteps steps of every scenario.  use_graph != 0 (and a non-null stream)
 * captures a 6-step cycle in a HIP graph once and replays it.  dt_hist
 * (host, optional): nscen x nsteps, row-major.  steps_done (host, optional,
 * nscen): steps completed per scenario.  Returns HX_E_DT_COLLAPSE if any
 * scenario's dt collapsed (that scenario stops; the others continue).
 */
int hx_ensemble_run(hx_ensemble *ens, double cfl, int nsteps, int use_graph, double *dt_hist,
                    int32_t *steps_done, void *stream);

/*
 * Two-level coupling for one fine box at refinement ratio 2: the pieces of
 * the reference's AMR level coupling (src/state/AmrMeshState_Physics.cpp:50-147,
 * src/state/AmrMeshState_AMR.cpp:667-844) and of the AMReX operations it
 * calls (FillPatchTwoLevels, FluxRegister, average_down).  The fine box
 * covers coarse cells [ci0, ci0 + nx_f/2) x [cj0, cj0 + ny_f/2); the fine
 * patch stores 2 ghost layers (HX_BC_GHOST sides) and its terrain ghosts.
 *
 * hx_face_fluxes: effective fluxes at listed faces of a level, re-evaluated
 * with the advance's face functions from the state the advance read --
 * plain Roe D+ = F + S+ (side 0, the low cell's view) or D- = F - S-
 * (side 1), the reference's side selection (src/solvers/Roe.cpp:440-575);
 * hydrostatic 0.5 (FL + FR) (src/solvers/HydrostaticRoe.cpp:612-627).
 * faces (device): nfaces x {dir, i, j, side}, x face (i, j) between cells
 * (i-1, j) and (i, j), 0 <= i <= nx, 0 <= j < ny (y faces alike; a face
 * outside that range yields NaN).  out (device, level dtype): nfaces/pair
 * rows of 3 values; row q receives scale F of faces q*pair .. q*pair+pair-1
 * (scale_x on x faces, scale_y on y faces), stored (accumulate 0) or added
 * one face at a time (accumulate 1).  With scales -dt dy, -dt dx on the
 * coarse faces (pair 1, FluxRegister::CrseInit COPY) and dt_f dy_f,
 * dt_f dx_f on the fine faces (pair 2, accumulate, FineAdd) this builds a
 * flux register.
 */
int hx_face_fluxes(hx_ctx *ctx, int solver, const hx_patch *level, int nfaces,
                   const int32_t *faces, int pair, double scale_x, double scale_y,
                   int accumulate, void *out, void *stream);

/*
 * Fine ghost cells (whole = 0) or the whole fine box with its ghosts
 * (whole = 1, the initial fill) from the coarse level at coarse time blend
 * alpha in [0, 1] (state (1 - alpha) crse_old + alpha crse_new; exactly the
 * old or new state at 0 and 1): eta = h + z and the momenta by conservative
 * linear interpolation with monotonized-central limited slopes (amrex
 * cell_cons_interp), then h = max(0, eta - z) on the fine cell's own bed
 * (fine->z, ghosts included).  h, hu, hv: the fine state's valid (0,0)
 * elements (writable, fine->pitch).
 */
int hx_fill_fine_ghosts(hx_ctx *ctx, const hx_patch *crse_old, const hx_patch *crse_new,
                        double alpha, int ratio, int ci0, int cj0, const hx_patch *fine,
                        int whole, void *h, void *hu, void *hv, void *stream);

/*
 * Reflux (AMReX FluxRegister::Reflux): reg (device) holds one register per
 * coarse face around the box, in the order x-low (ncy faces), x-high (ncy),
 * y-low (ncx), y-high (ncx); the coarse cell outside each face gains
 * -reg/(dx dy) on the low sides and +reg/(dx dy) on the high sides.
 */
int hx_reflux(hx_ctx *ctx, const hx_patch *crse, void *h, void *hu, void *hv, int ci0, int cj0,
              int ncx, int ncy, const void *reg, void *stream);

/*
 * Covered coarse cells from the fine valid cells: the 2 x 2 children summed
 * j-outer, i-inner, times 1/4 (amrex::average_down of the state; the
 * reference's AverageDownTo, src/state/AmrMeshState_AMR.cpp:836-844).
 */
int hx_average_down_state(hx_ctx *ctx, const hx_patch *fine, int ratio, int ci0, int cj0,
                          const hx_patch *crse, void *h, void *hu, void *hv, void *stream);

/*
 * Multi-box hierarchy pieces (levels held as full-domain arrays with two
 * cells of padding, each level a list of boxes; hydroexa_amd/amr.py's
 * AmrHierarchy drives them).
 *
 * hx_fill_fine_region: hx_fill_fine_ghosts for a box of a level that other
 * boxes share: cells whose skip byte is non-zero (skip: the fine level's
 * valid-cell mask at the box's (0,0), skip_pitch bytes between rows, its
 * ghost ring addressable) are left as they are.  whole = 1 fills the box and
 * its ring (a level remade after regridding: skip = the cells valid before).
 * The box may touch the domain boundary; ring cells outside the domain are
 * written into the padding and never read.
 */
int hx_fill_fine_region(hx_ctx *ctx, const hx_patch *crse_old, const hx_patch *crse_new,
                        double alpha, int ratio, int ci0, int cj0, const hx_patch *fine, int whole,
                        const uint8_t *skip, int64_t skip_pitch, void *h, void *hu, void *hv,
                        void *stream);

/*
 * hx_reflux for one side of a box (orient 0 x-low, 1 y-low, 2 x-high,
 * 3 y-high, AMReX's Orientation order; -1 all sides).  Refluxing every box
 * side by side in that order reproduces FluxRegister::Reflux's order of
 * additions at a coarse cell two boxes border.  The box may touch the
 * domain boundary: faces there have no coarse cell outside and are skipped.
 */
int hx_reflux_side(hx_ctx *ctx, const hx_patch *crse, void *h, void *hu, void *hv, int ci0,
                   int cj0, int ncx, int ncy, const void *reg, int orient, void *stream);

/*
 * Regrid tag blocks (device in and out): blocks[(j/b) * (nx/b) + i/b] = 1 when
 * a valid cell of the b x b block lies within nbuf cells (max norm) of a
 * cell that is tagged and valid; tags: hx_tag_level's bytes (tag_pitch per
 * row), valid: the level's mask at its (0,0) (valid_pitch per row).  The
 * n_error_buf growth and blocking-factor coarsening of AMReX's
 * MakeNewGrids in one pass.
 */
int hx_tag_blocks(hx_ctx *ctx, const uint8_t *tags, int64_t tag_pitch, const uint8_t *valid,
                  int64_t valid_pitch, int nx, int ny, int nbuf, int b, uint8_t *blocks,
                  void *stream);

/*
 * New grids of level k + 1 from level k's tag blocks (host in and out; no
 * device work): Berger-Rigoutsos clustering of the blocks that are tagged
 * and allowed (the proper-nesting domain) at efficiency grid_eff, clusters
 * reaching outside the allowed blocks replaced by the allowed rectangles
 * inside them that hold tags, refined by b * ratio and chopped to
 * max_grid_size.  tags, allowed: nby x nbx bytes, row-major.  out receives
 * *nout boxes {i0, j0, w, h} in level k + 1's cells, sorted by (j0, i0, w,
 * h); HX_E_BADARG with *nout set when cap is too small.  No counterpart in
 * the reference, which calls AMReX's AmrMesh::MakeNewGrids
 * (src/state/AmrMeshState_Physics.cpp:50-105); this is that grid
 * generation's clustering and chopping, done natively.
 */
int hx_cluster_blocks(const uint8_t *tags, const uint8_t *allowed, int nby, int nbx, int b,
                      int ratio, double grid_eff, int max_grid_size, int32_t *out, int cap,
                      int32_t *nout);

/*
 * hx_level_dt over the valid cells of nbox boxes of one level (host array
 * of patches), one reduction and one synchronisation.
 */
int hx_boxes_dt(hx_ctx *ctx, int solver, int nbox, const hx_patch *boxes, double cfl,
                double *dt_out, double *max_vel_out, void *stream);

/*
 * A level of an AMR hierarchy held whole in device memory (hx_amr_run).
 * Arrays are (ny + 4) x (nx + 4) per component with two cells of padding,
 * valid cell (0,0) at [2][2], pitch nx + 4; state buffers hold h, hu, hv
 * planes, terrain z, n planes.  patches: rectangles (even origin and size)
 * covering exactly the cells the mask marks; regs: the level's flux
 * registers as a fine level, 6 (w/2 + h/2) values per patch in patch order.
 */
#define HX_AMR_MAX_LEVELS 8
typedef struct hx_amr_level {
    int32_t nx, ny;
    int32_t npatch;
    int32_t cur;                  /* buffer holding the current state (in/out) */
    const int32_t *patches;       /* host, npatch x {i0, j0, w, h}              */
    void *u[2];                   /* device state buffers                        */
    const void *terrain;          /* device                                      */
    const uint8_t *mask;          /* device, 1 on the cells the level owns       */
    void *regs;                   /* device                                      */
    double dx, dy;
    double t_new, t_old;          /* clocks (in/out)                             */
    int32_t istep, last_regrid;   /* (in/out)                                    */
} hx_amr_level;

/* Regrid request from hx_amr_run: level lev's grids are due at time; the
 * host may change levels lev + 1 .. (patches, masks, registers, new levels
 * appended, *nlevels) in the arrays it passed.  Non-zero aborts the run. */
typedef int (*hx_amr_regrid_fn)(void *user, int lev, double time);

/*
 * Coarse steps of an AMR hierarchy until *time reaches t_end or max_steps
 * steps, launched from C++: per step ComputeDt (BaseSolver::compute_dt at
 * ratio 2, lockstep sharing the finest dt; src/state/AmrMeshState_Physics.cpp:13-45)
 * with one synchronisation, then TimeStepWithSubcycling (:50-147): regrid
 * requests every regrid_int steps of a level, ghost rings from the coarser
 * level at the substep's time (clock-ratio blend), side-selected flux
 * registers (faces generated on the device), the fused advance per patch,
 * reflux in AMReX's side order and average_down -- the same kernels and
 * order as hydroexa_amd.amr.AmrHierarchy.step.  dt_hist (host, optional):
 * the coarse dt of each step.  Returns HX_E_DT_COLLAPSE when a level's
 * estimate falls below 1e-12 (before that step).
 */
int hx_amr_run(hx_ctx *ctx, int solver, int dtype, hx_amr_level *levels, int *nlevels,
               int max_level, const int32_t *bc_lo, const int32_t *bc_hi, int subcycle,
               int do_reflux, int regrid_int, hx_amr_regrid_fn regrid, void *user, double cfl,
               double t_end, int max_steps, double *time, double *dt_hist, int32_t *steps_done,
               void *stream);

/*
 * nsteps coupled steps of one coarse level and one fine box at ratio 2,
 * without host synchronisation: the reference's recursive subcycled step
 * (AmrMeshState::TimeStepWithSubcycling, src/state/AmrMeshState_Physics.cpp:50-147)
 * with ComputeDt (:13-45) for two levels, built from the four calls above
 * and the fused advance.  Per step: dt from both levels' CFL by the
 * multi-level rule (lockstep: one shared dt), coarse register + advance,
 * 2 fine substeps (1 in lockstep) of ghost fill + register + advance,
 * reflux, average_down.  dt is computed on the device; the coarse advance
 * reduces the CFL of its output outside the box grown by one cell, which a
 * separate pass reduces after reflux and average_down.
 *
 * crse: coarse state (physical sides); ch/chu/chv its alternate buffer
 * (crse pitch).  fine: fine box patch (ng >= 2, HX_BC_GHOST on every side,
 * terrain with its ghosts, dx = crse dx / 2), covering coarse cells
 * [ci0, ci0 + nx/2) x [cj0, cj0 + ny/2), one coarse cell away from every
 * domain edge; fh/fhu/fhv its alternate buffer (valid (0,0), fine pitch,
 * ghost storage included).  reg (device): 6 (ncx + ncy) values of the level
 * dtype.  dt_hist (host, optional): 2 x nsteps (coarse dt, fine dt).
 * After the call the coarse state is in crse's buffer when *steps_done is
 * even, else in ch/chu/chv; the fine state is in fine's buffer when the
 * number of fine substeps done (2 per step subcycled, 1 lockstep) is even.
 * Returns HX_E_DT_COLLAPSE (with *steps_done) when either level's estimate
 * falls below 1e-12.  Synchronises the stream at the end.
 */
int hx_two_level_run(hx_ctx *ctx, int solver, const hx_patch *crse, void *ch, void *chu,
                     void *chv, const hx_patch *fine, void *fh, void *fhu, void *fhv, int ci0,
                     int cj0, int subcycle, double cfl, int nsteps, void *reg, double *dt_hist,
                     int32_t *steps_done, void *stream);

/*
 * hx_two_level_run's nsteps coupled steps (1 <= nsteps <= 4096) as stream
 * work only: kernel launches on stream, no host copy, no allocation and no
 * synchronisation, so that a caller may capture the call in a HIP graph and
 * replay it.  No reference counterpart: it is for callers that capture the
 * coupled step (graph replay of a fixed-size coarse step loop).  Arguments
 * as for hx_two_level_run; each call starts from the CFL of its inputs.
 * dt_hist_dev (device, optional): receives 2 x nsteps (coarse dt, fine dt);
 * when null they stay in the context.  status_dev (device, optional): one
 * int32 set at the end of the work to 0, or to k + 1 when the estimate of
 * step k fell below 1e-12 (the steps from k on then leave the state as it
 * was).  The final state lies where hx_two_level_run leaves it after
 * nsteps steps: with nsteps even, in crse's and fine's own buffers.
 * Calls on one context share its work area: run them in stream order.
 */
int hx_two_level_enqueue(hx_ctx *ctx, int solver, const hx_patch *crse, void *ch, void *chu,
                         void *chv, const hx_patch *fine, void *fh, void *fhu, void *fhv,
                         int ci0, int cj0, int subcycle, double cfl, int nsteps, void *reg,
                         double *dt_hist_dev, int32_t *status_dev, void *stream);

/*
 * Adjoint of a single-level rollout (FP64).  No reference counterpart:
 * HydroEXA computes no derivatives.  The rollout is nsteps of the step
 * behind hx_advance_level (plain Roe or hydrostatic Roe with Manning
 * friction, src/solvers/Roe.cpp:250-434, src/solvers/HydrostaticRoe.cpp:400-606)
 * with physical sides; its outputs are the depth at ngauge cells after
 * every step (the gauge series) and the final state.  hx_adjoint_forward
 * runs the rollout and keeps the state every ckpt_every steps;
 * hx_adjoint_backward recomputes each segment from its checkpoint and runs
 * the hand-written adjoint step backwards through it, with the dt schedule
 * of the forward held fixed.
 *
 * geom: nx, ny, dx, dy and bc_lo/bc_hi (HX_BC_REFLECT / HX_BC_TRANSMIT),
 * dtype HX_F64; its pointers are ignored.  gauges (host): flat cell
 * indices j*nx + i.  Device memory: ncp + max(ckpt_every-1, 2) + 2 states
 * of 3*nx*ny doubles (+1 for the hydrostatic solver), ncp =
 * ceil(nsteps / ckpt_every).
 */
typedef struct hx_adjoint hx_adjoint;
int hx_adjoint_create(hx_ctx *ctx, int solver, const hx_patch *geom, int nsteps, int ckpt_every,
                      int ngauge, const int64_t *gauges, hx_adjoint **out);
int hx_adjoint_destroy(hx_adjoint *adj);

/*
 * Forward rollout.  state: t^0 (dense, pitch nx, F64, no eta) with z and n,
 * which must stay alive until the backward pass is done.  cfl > 0: each
 * step's dt is cfl times the CFL estimate of its input (as hx_run_level),
 * and dt_sched (host, nsteps) receives the schedule; cfl <= 0: dt_sched
 * gives it.  series (device, nsteps x ngauge): depth at the gauges after
 * each step.  h_out/hu_out/hv_out (device, dense): the final state.
 */
int hx_adjoint_forward(hx_adjoint *adj, const hx_patch *state, double cfl, double *dt_sched,
                       double *series, void *h_out, void *hu_out, void *hv_out,
                       int32_t *steps_done, void *stream);

/*
 * Backward pass of the last forward.  g_series (device, nsteps x ngauge,
 * optional) and g_final (device, 3 planes h, hu, hv, optional): adjoints of
 * the outputs.  g_state (device, 3 planes): gradient with respect to t^0;
 * g_z, g_n (device, nx*ny): gradients with respect to bed and roughness.
 */
int hx_adjoint_backward(hx_adjoint *adj, const double *g_series, const double *g_final,
                        double *g_state, double *g_z, double *g_n, void *stream);

/*
 * Batched forward + adjoint over nscen independent rollouts (flood
 * scenarios of uneven sizes): one launch per forward step, per recomputed
 * step and per adjoint kernel covers every scenario, each with its own
 * geometry, terrain, gauges, dt schedule and checkpoints.  No reference
 * counterpart (the reference neither batches nor differentiates).  All
 * scenarios share the solver, nsteps and ckpt_every, and the results of a
 * scenario equal the one-scenario object's (hx_adjoint_create is the
 * nscen = 1 case) bit for bit.
 *
 * create: geoms[nscen] as hx_adjoint_create's geom; ngauge[nscen] and
 * gauges[nscen] (host arrays of flat cell indices).
 * forward: states[nscen] as hx_adjoint_forward's state; dt_sched (host,
 * nscen x nsteps, scenario-major) receives (cfl > 0) or gives (cfl <= 0)
 * the schedules; series[nscen] (device, nsteps x ngauge[s]; entries of
 * scenarios without gauges may be null); finals[3 nscen] (device planes
 * h, hu, hv per scenario); steps_done[nscen] (host, optional).  A scenario
 * whose dt collapses stops and HX_E_DT_COLLAPSE is returned (no backward).
 * backward: per-scenario arrays of device pointers; g_series and g_final
 * (the arrays, or their entries) may be null.
 */
int hx_adjoint_batch_create(hx_ctx *ctx, int solver, int nscen, const hx_patch *geoms, int nsteps,
                            int ckpt_every, const int *ngauge, const int64_t *const *gauges,
                            hx_adjoint **out);
int hx_adjoint_batch_forward(hx_adjoint *adj, const hx_patch *states, double cfl,
                             double *dt_sched, double *const *series, void *const *finals,
                             int32_t *steps_done, void *stream);
int hx_adjoint_batch_backward(hx_adjoint *adj, const double *const *g_series,
                              const double *const *g_final, double *const *g_state,
                              double *const *g_z, double *const *g_n, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HYDROEXA_H */
