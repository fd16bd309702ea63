/*
 * fomo_hip.h — C ABI of the MI355X render engine behind the FoMo-C API.
 *
 * Plain C: pointers, sizes and int status codes; no C++ types or exceptions cross it.
 * Every function returns FOMO_OK (0) or a negative FOMO_ERR_* code; the message of the
 * last failure on the calling thread is returned by fomo_last_error().
 *
 * Device state (datasets, emission cubes, G(T,n) tables, per-view scratch) is resident in
 * HBM and addressed by integer ids. Host buffers passed in are read or written
 * synchronously before the call returns; the caller owns them.
 *
 * Citations name the reference function each entry point replaces (fomo-c/src/...).
 */
#ifndef FOMO_HIP_H
#define FOMO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FOMO_OK 0
#define FOMO_ERR_ARG (-1)      /* invalid argument or size */
#define FOMO_ERR_HIP (-2)      /* HIP runtime failure */
#define FOMO_ERR_NOTFOUND (-3) /* unknown id */
#define FOMO_ERR_NOMEM (-4)    /* device allocation failed or over budget */
#define FOMO_ERR_STATE (-5)    /* not initialised / wrong object kind */
#define FOMO_ERR_IO (-6)       /* file missing or malformed */
#define FOMO_ERR_CANCELLED (-7) /* server request cancelled before it ran */

/* Message for the last failing call on this thread ("" if none). */
const char* fomo_last_error(void);

/* Library version string. */
const char* fomo_version(void);

/* Select and initialise a device (0 = first visible GPU). Idempotent. */
int fomo_init(int device);
/* Release every resident object and the device context. */
int fomo_shutdown(void);
/* Device name, HBM bytes and compute-unit count of the initialised device. */
int fomo_device_info(char* name, int name_len, double* hbm_bytes, int* compute_units);

/* ---------------------------------------------------------------------------------------
 * Datasets. Replaces DataCube::setdata / push_back (fomo-datacube.cpp:205-279) on the device
 * side: SoA float32 coordinates (dim 2 or 3; z is taken as 0 when dim==2) and nvars float32
 * variables per point. The upload builds the view-invariant spatial index (uniform-grid cell
 * list from a radix sort) that replaces the Boost R-tree (fomo-nearestneighbour.cpp:93-100).
 * cell_scale multiplies the default cell edge (V/N)^(1/3); 0 selects the default 1.0.
 * ------------------------------------------------------------------------------------- */
int fomo_dataset_create(int64_t n, int dim, const float* x, const float* y, const float* z,
                        int nvars, const float* const* vars, double cell_scale, int* out_id);
/* Streaming ingest for loaders (no counterpart in the reference, where a loader fills a DataCube
 * point by point with push_back_datapoint or whole vectors with setdata, fomo-datacube.cpp:205-279,
 * and the render copies it again): fomo_dataset_begin creates an empty dataset of n points on the
 * device, fomo_dataset_append copies a chunk of `count` points starting at point `offset` (SoA
 * float32, z ignored when dim == 2) straight to device memory before it returns, in any order and
 * size, and fomo_dataset_commit builds the spatial index once every point has been written
 * exactly once (FOMO_ERR_STATE otherwise). Until the commit the id is known only to these calls
 * and fomo_dataset_destroy; afterwards the dataset is the one fomo_dataset_create makes from the
 * same points. */
int fomo_dataset_begin(int64_t n, int dim, int nvars, int* out_id);
int fomo_dataset_append(int id, int64_t offset, int64_t count, const float* x, const float* y, const float* z,
                        const float* const* vars);
int fomo_dataset_commit(int id, double cell_scale);
int fomo_dataset_destroy(int id);
/* n points, cells per axis, maximum points in one cell. */
int fomo_dataset_info(int id, int64_t* n, int* grid_dims3, int* max_cell_count);
/* Rebuild the spatial index with another cell_scale (used by the autotuner). */
int fomo_dataset_reindex(int id, double cell_scale);
/* Candidate lookup grid (the NN fast path): subdiv lookup cells per cell-list edge (default 2);
 * 0 disables it, leaving the exact ring search over the cell list. Results are identical either
 * way; only speed and memory change. */
int fomo_dataset_set_lookup(int id, int subdiv);
/* Ray tile of the lookup pass's blocks (k_nn): width 8 (default), 16, 32 or 64 rays, 256 rays per
 * block (the exact-query pass always takes 4 x 4 rays through 4 LOS planes per wave). Speed only;
 * results do not change. */
int fomo_dataset_set_ray_tile(int id, int width);
/* Node steps an exact octree query of the NN passes takes (per wave of 64 samples) before its
 * samples move to a pass that walks 1024 node ranges of the tree with 32 waves in parallel and
 * merges them (default 2048; 0: never). Speed only; results do not change. */
int fomo_dataset_set_far_budget(int id, int steps);
/* Tetra path counts since the triangulation was built: out2[0] = LOS samples whose filtered FP64
 * walk could not place them (an orientation test within its error bound at the tet where the walk
 * ended, or the walk's step cap) and that were walked again with exact orientation tests;
 * out2[1] = samples left unresolved (beyond a view's redo list of 65536, or at the exact walk's
 * step cap): those keep the filtered walk's tet, or none at the cap. Both 0 before the first tetra
 * render. Reads a snapshot while other streams may still render. */
int fomo_dataset_tet_stats(int id, int64_t* out2);
/* Active subdivision (0 = disabled), lookup cells, candidates summed over the cells (a far group's
   shared list counts once per cell that uses it), cells left to the exact octree query. */
int fomo_dataset_lookup_info(int id, int* subdiv, int64_t* cells, int64_t* candidates, int64_t* fallback_cells);

/* Rectilinear datasets. When the points form a product grid (x, y and z each take a set of
 * distinct values, and every combination occurs exactly once: structured simulation output), the
 * index is the three sorted axes instead of the cell list and candidate lookup grid. A target's
 * nearest point then combines the axis values that bracket it (exact ties are resolved as by every
 * other query: squared float-coordinate distances in FP64, then the lowest original index), and
 * the exact octree query answers the rest. Detection is on by default; on = 0 re-indexes the
 * dataset as scattered points, on = 1 detects again. Results are identical either way; speed and
 * memory change (no 128-byte lookup record per sample, no lookup grid). Replaces, like the rest of
 * the index, the per-view R-tree build of fomo-nearestneighbour.cpp:93-100. */
int fomo_dataset_set_rectilinear(int id, int on);
/* Values per axis of a rectilinear dataset's grid (axes3), all 0 for a scattered one. */
int fomo_dataset_rect_info(int id, int* axes3);
/* Index kind: 0 scattered points (cell list and candidate lookup grid), 1 rectilinear (a full
 * product grid, indexed by its axes), 2 masked grid (the points fill at least a quarter of a
 * product grid's nodes, one point per node, as a Cartesian mesh cut to a region: the candidate
 * lookup grid plus the grid's axes and a node -> point map, which the NN queries of the
 * unfused passes try first; a target whose nearest node holds no point takes the lookup grid).
 * axes3: values per axis of the grid (0 for scattered points). */
int fomo_dataset_grid_info(int id, int* kind, int* axes3);

/* Far and outer groups of the lookup grid (targets more than 2 cells from any point, and targets
 * beyond the grid, whose cells read a group's shared candidate list). counts5: far groups, far
 * groups with a list, far groups whose cells all take their list (group headers), outer groups,
 * outer groups with a list. All 0 when the dataset has none. */
int fomo_dataset_lookup_groups(int id, int64_t* counts5);

/* ---------------------------------------------------------------------------------------
 * G(T,n) tables. The host-side parse of readgoftfromchianti (fomo-CHIANTI.cpp:143-218) is
 * fomo_table_read; fomo_table_create uploads a parsed table (nt log T values, nrho log n
 * values, nrho*nt G values with T fastest).
 * ------------------------------------------------------------------------------------- */
/* Parse a table file. With logt==NULL only the header and sizes are filled. */
int fomo_table_read(const char* path, char* ion, int ion_len, double* lambda0, double* atweight,
                    char* unit, int unit_len, int* nt, int* nrho, float* logt, float* logn, float* g);
int fomo_table_create(int nt, int nrho, const float* logt, const float* logn, const float* g, int* out_id);
int fomo_table_destroy(int id);

/* ---------------------------------------------------------------------------------------
 * Emission cubes (the GoftCube on the device).
 * fomo_emission_chianti replaces emissionfromdatacube (fomo-CHIANTI.cpp:228-322): variables 0,1
 * of the dataset are n and T, 2..4 the velocity. Output per point: peak (n^2 G / fwhm times the
 * normalisation) and fwhm, computed in FP64 and stored float32. imaging!=0 selects the Imaging
 * branch (fwhm=1, normalisation 1); abundratio scales the spectroscopic normalisation.
 * fomo_emission_thomson replaces thomsonfromdatacube (fomo-thomson.cpp:60-92): n*tot and n*pol.
 * ------------------------------------------------------------------------------------- */
int fomo_emission_chianti(int dataset, int table, double lambda0, double atweight, int imaging,
                          double abundratio, int* out_goft);
/* fomo_emission_chianti with the G(T, n) interpolation chosen: interp 0 is the reference's (linear
 * in log T, nearest density row, fomo-CHIANTI.cpp:64-121, what fomo_emission_chianti does); interp
 * 1, an extension with no counterpart in the reference, is bilinear in (log T, log n), zero outside
 * the table's cells. */
int fomo_emission_chianti_interp(int dataset, int table, double lambda0, double atweight, int imaging,
                                 double abundratio, int interp, int* out_goft);
int fomo_emission_thomson(int dataset, int* out_goft);
/* Emission cube of a dataset that holds a GoftCube already computed on the host: variables 0 and
 * 1 are (peak, fwhm) of a line at lambda0 (kind 0; variables 2..4 the velocity), or (n tot,
 * n pol) for Thomson scattering (kind 1). This is the input of the reference's RenderWith*
 * functions (FoMo-internal.h:14-30, e.g. RenderWithNearestNeighbour, fomo-nearestneighbour.cpp:
 * 284-306); the renders use the given values as emissionfromdatacube / thomsonfromdatacube would
 * have produced them. A cube of this kind cannot share a batch with cubes computed from T. */
int fomo_emission_given(int dataset, int kind, double lambda0, int* out_goft);
int fomo_goft_destroy(int id);
/* Copy the two per-point emission variables back, in the dataset's original point order. */
int fomo_goft_read(int goft, float* var0, float* var1);

/* ---------------------------------------------------------------------------------------
 * Rendering. meta (optional) receives 10 doubles per view:
 *   minx,maxx,miny,maxy,minz,maxz (rotated extents), maxdistance, pathlength, scale, apix.
 * Output intensity layout is RenderCube's: index (iy*nx + jx)*nl + il, lambda fastest.
 * ------------------------------------------------------------------------------------- */

/* One view of nearestneighbourinterpolation (fomo-nearestneighbour.cpp:31-281).
 * lambda_width in m/s; dn!=0 applies the DN/arcsec scaling (:252-265). out: ny*nx*max(nl,1). */
int fomo_render_nn(int goft, double l, double b, int nx, int ny, int nz, int nl, double lambda_width,
                   int dn, float* out, double* meta);

/* One view of thomsoninterpolation (fomo-thomson.cpp:94-367): pB image, out: ny*nx. */
int fomo_render_thomson(int goft, double l, double b, int nx, int ny, int nz, float* out, double* meta);

/* One view of CGALinterpolation (fomo-CGAL.cpp:75-307): each LOS sample takes the values of the
 * nearest vertex of its Delaunay tet, 0 outside the convex hull (:194-228); samples at
 * z = k/(nz-1)*(maxz-minz)+minz (:189). linear != 0 (an extension, method "Linear") interpolates
 * the tet's vertex values barycentrically instead. The dataset is triangulated on first use (host
 * Delaunay builder, replacing triangulationfromdatacube, :33-73). Arguments as fomo_render_nn. */
int fomo_render_tet(int goft, double l, double b, int nx, int ny, int nz, int nl, double lambda_width,
                    int dn, int linear, float* out, double* meta);

/* One view of projectioninterpolation (fomo-projection.cpp:17-214), method "Projection": every
 * point adds its line profile (nl > 1) or its peak (nl == 1) to the pixel its rotated (x', y')
 * rounds to (:126-158); no LOS samples. nz only sets the path length (:163-170). Each pixel sums
 * its points in original point order in float, the sums of a one-thread run of the reference
 * (which adds with OpenMP atomics, in no fixed order). Arguments and output as fomo_render_nn;
 * meta[6] (the NN search half-width) has no meaning here. */
int fomo_render_projection(int goft, double l, double b, int nx, int ny, int nz, int nl,
                           double lambda_width, int dn, float* out, double* meta);

/* Allocation-free variant of fomo_render_nn. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc). It runs the same kernels as fomo_render_nn and gives the same image
 * bit for bit, but takes its per-view buffers from one static device array of
 * FOMO_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) and calls no hipMalloc,
 * hipFree or allocating library routine. Views whose buffers exceed the array fail with
 * FOMO_ERR_ARG: a single-line view of at least 65,536 rays needs about 5 bytes per ray, any
 * other view about 16 bytes per LOS sample. Kernel profiling (fomo_profile_enable) must be off.
 * out_dev: device buffer of ny*nx*max(nl,1) floats (RenderCube layout); meta_dev (optional):
 * device buffer of the 10 meta doubles. The work is enqueued on `stream` (a hipStream_t; NULL =
 * the library's stream) and the call returns without waiting: synchronize that stream before
 * reading out_dev. Every call on one GPU shares its array, so calls whose work overlaps on the
 * device race; serialize them on one stream. The emission cube and its dataset must outlive the
 * enqueued work. */
#define FOMO_NOALLOC_SCRATCH_BYTES (64ll << 20)
int fomo_render_nn_noalloc(int goft, double l, double b, int nx, int ny, int nz, int nl, double lambda_width,
                           int dn, float* out_dev, double* meta_dev, void* stream);

/* The projection's pixel iy*nx + jx of every point (original order) for one view. For parity
 * checks of the pixel assignment (:129-131). */
int fomo_projection_pixels(int dataset, double l, double b, int nx, int ny, int nz, int32_t* out_pixel,
                           double* meta);

/* Autotuner (no counterpart in the reference): for the dataset of the ngoft emission cubes (one
 * dataset, one kind; the lines of a batch), times the nearest-neighbour passes (k_nn + k_nn_far)
 * of one representative view (l, b, nx x ny x nz; Thomson geometry for a Thomson cube), gathering
 * the same per-point records a batch of these cubes gathers, over cell-list scales {0.7, 1, 1.4} x
 * lookup subdivisions {1, 2, 3} x ray tiles of the lookup pass's blocks {8x32, 16x16, 32x8, 64x4},
 * `reps` timed repetitions each. The dataset keeps the fastest configuration when it beats the
 * one on entry by more than 3 %, else the one on entry. result (optional, 5 doubles): chosen cell
 * scale, subdivision, its ms per view, the ms per view on entry, the chosen tile width in rays.
 * Render results do not depend on the configuration (NN choices are exact for every one). */
int fomo_autotune(int ngoft, const int* gofts, double l, double b, int nx, int ny, int nz, int reps,
                  double* result);

/* Build (once; again after fomo_dataset_reindex) the Delaunay triangulation of a dataset and
 * upload it. ntets counts finite and infinite tets; seconds is the build time. */
int fomo_dataset_triangulate(int dataset, int64_t* ntets, double* seconds);

/* The tetra path's vertex choice for every LOS sample of one view (original point indices, -1 =
 * outside the convex hull), layout (iy*nx + jx)*nz + k. For parity checks. meta: 12 doubles, the
 * 10 of a render, then this view's samples walked again with exact orientation tests and those
 * left unresolved (see fomo_dataset_tet_stats). */
int fomo_tet_vertices(int dataset, double l, double b, int nx, int ny, int nz, int32_t* out_vertex,
                      double* meta);

/* Nearest-neighbour choice of every LOS sample of one view (original point indices, -1 = none),
 * layout (iy*nx + jx)*nz + k. thomson!=0 uses the Thomson geometry (fomo-thomson.cpp:136-317).
 * For parity checks of the spatial index against the reference's R-tree query. meta: 12 doubles,
 * the 10 of a render, then the waves of 64 samples whose exact query reached the far budget
 * (fomo_dataset_set_far_budget) and the number of them the parallel range walk takes (the rest
 * finish their own walk). */
int fomo_nn_indices(int dataset, double l, double b, int nx, int ny, int nz, int thomson,
                    int32_t* out_idx, double* meta);

/* Batched render: every view x every emission cube (spectral lines) of one dataset, in one
 * captured HIP graph. The cell list is shared by all views, the NN result of a view by all lines.
 * No counterpart in the reference: it replaces the view loop of RenderWithNearestNeighbour
 * (fomo-nearestneighbour.cpp:285-306) run once per line. Results stay resident on the device;
 * copy them with fomo_batch_read. */
int fomo_batch_create(int ngoft, const int* gofts, int nviews, const double* l, const double* b,
                      int nx, int ny, int nz, int nl, double lambda_width, int dn, int* out_batch);
int fomo_batch_run(int batch);               /* launch the graph and wait */
int fomo_batch_launch(int batch);            /* launch the graph, do not wait */
int fomo_batch_sync(int batch);
/* Run the batch's kernels once eagerly (same kernels, same order, same stream) with HIP events
 * around each launch, accumulating per-kernel durations into the profile stats below. */
int fomo_batch_profile(int batch);
/* out: [view][goft][ny][nx][nl] floats; meta: nviews*10 doubles (either may be NULL). */
int fomo_batch_read(int batch, float* out, double* meta);
/* One image of a batch: view `view`, emission cube number `line` of the batch (ny*nx*max(nl,1)
 * floats), and that view's 10 meta doubles (either may be NULL). */
int fomo_batch_read_view(int batch, int view, int line, float* out, double* meta);
int fomo_batch_destroy(int batch);

/* ---------------------------------------------------------------------------------------
 * Render server: continuous batching of independent view requests over resident datasets and
 * emission cubes. No counterpart in the reference (its render() is one synchronous call per
 * object, fomo-object.cpp:374-440); each request computes exactly what fomo_render_nn /
 * fomo_render_thomson compute for the same arguments.
 *
 * A worker thread admits queued requests in arrival order whenever one of max_inflight slots and
 * a block of the server's device arena (device_bytes: per-sample buffers and image of a request,
 * see fomo_server_footprint) and pinned host arena (host_bytes: image staging) are free; smaller
 * requests may pass one that does not fit, a bounded number of times. Admitted requests run on
 * nstreams HIP streams (default priority) and overlap on the device. The image is copied to
 * `out` (ny*nx*max(nl,1) floats, RenderCube layout) and `meta` (10 doubles, as above, or NULL)
 * before the ticket turns done; both must stay valid until then. Ticket states:
 * 0 queued, 1 running, 2 done, negative = failed with that FOMO_ERR_* code.
 * Datasets and emission cubes referenced by queued or running requests cannot be destroyed or
 * re-indexed (FOMO_ERR_STATE). Destroying a server fails its queued requests and waits for the
 * running ones.
 * ------------------------------------------------------------------------------------- */
#define FOMO_METHOD_NN 0      /* "NearestNeighbour": fomo_render_nn */
#define FOMO_METHOD_THOMSON 1 /* "Thomson": fomo_render_thomson (nl, lambda_width, dn ignored) */
#define FOMO_METHOD_TET 2     /* "CGAL": fomo_render_tet (dataset must be triangulated first) */
#define FOMO_METHOD_LINEAR 3  /* "Linear": fomo_render_tet with linear != 0 (likewise) */
#define FOMO_METHOD_PROJECTION 4 /* "Projection": fomo_render_projection */
int fomo_server_create(int64_t device_bytes, int64_t host_bytes, int max_inflight, int nstreams,
                       int* out_server);
int fomo_server_destroy(int server);
/* Memory a request of this shape takes from the device and host arenas while it runs. */
int fomo_server_footprint(int goft, int method, int nx, int ny, int nz, int nl,
                          int64_t* device_bytes, int64_t* host_bytes);
/* Queue one view; FOMO_ERR_NOMEM when its footprint exceeds the server's arenas. */
int fomo_server_submit(int server, int goft, int method, double l, double b, int nx, int ny,
                       int nz, int nl, double lambda_width, int dn, float* out, double* meta,
                       int64_t* out_ticket);
int fomo_server_poll(int server, int64_t ticket, int* state);
/* Block until the ticket is done or failed (timeout_s < 0: no limit); FOMO_ERR_STATE on timeout,
 * the ticket's own error code when it failed. */
int fomo_server_wait(int server, int64_t ticket, double timeout_s, int* state);
int fomo_server_wait_all(int server, double timeout_s);
/* Cancel a queued request: it leaves the queue and its ticket fails with FOMO_ERR_CANCELLED; its
 * out and meta are not written. A request already admitted (running) or finished cannot be
 * cancelled (FOMO_ERR_STATE). Cancelled requests count as failed in fomo_server_stats. */
int fomo_server_cancel(int server, int64_t ticket);
/* Seconds since the server was created: submitted, admitted (-1 if queued), finished (-1). */
int fomo_server_times(int server, int64_t ticket, double* times3);
/* 12 doubles: submitted, completed, failed, queued, running, peak running, device bytes in use,
 * device peak, device budget, host bytes in use, host peak, host budget. */
int fomo_server_stats(int server, double* stats12);

/* ---------------------------------------------------------------------------------------
 * Kernel timing with HIP events on the library's launch stream (for roofline reporting).
 * When enabled, each launch of a named kernel is bracketed by events; stats return the
 * number of timed launches and their total milliseconds.
 * ------------------------------------------------------------------------------------- */
int fomo_profile_enable(int on);
int fomo_profile_reset(void);
int fomo_profile_stats(const char* kernel, int64_t* count, double* total_ms);
/* Device bandwidth probes (no counterpart in the reference). fomo_copy_probe copies `bytes`
 * between two device buffers `reps` times with a one-pass nontemporal float4 streaming kernel
 * (k_copy_f4); mean milliseconds per copy (a copy of B bytes moves B read + B written; also the
 * calibration of hardware byte counters). fomo_gather_probe reads `lines` random aligned 128-byte
 * lines of a buffer of `bytes` bytes, `reps` times (k_gather_lines, the access pattern of the
 * NN passes' record gathers); mean milliseconds per launch. */
int fomo_copy_probe(int64_t bytes, int reps, double* ms_per_copy);
int fomo_gather_probe(int64_t bytes, int64_t lines, int reps, double* ms_per_gather);

/* ---------------------------------------------------------------------------------------
 * Synthetic inputs (host-only, deterministic from the seed; no GPU needed).
 * fomo_gen_loop_cube: the straight cylindrical loop in a 10 Mm box on an m^3 grid
 *   (x,y,z in [-5,5] Mm; loop axis z, radius 1 Mm, tanh edge 0.1 Mm; n 1e9.5 inside / 1e9 out,
 *   T t_in / t_out; lognormal noise sigma on every field; Gaussian kink v of amplitude vamp m/s).
 *   jitter (cells) displaces points uniformly; sausage (relative amplitude, wavelength 5 Mm)
 *   modulates the density inside the loop. Outputs have m^3 entries each.
 * fomo_gen_goft_table: G = 1e-24 exp(-(logT-peak)^2/(2 sigma^2)) (1+0.05 (logn-10)) on
 *   log T in [4,8] (nt points) x log n in [8,12] (nrho points), written in the CHIANTI table
 *   text format read by fomo_table_read.
 * ------------------------------------------------------------------------------------- */
int fomo_gen_loop_cube(int m, uint64_t seed, double t_in, double t_out, double noise_sigma,
                       double vamp, double jitter, double sausage, float* x, float* y, float* z,
                       float* rho, float* T, float* vx, float* vy, float* vz);
int fomo_gen_goft_table(const char* path, const char* ion, double lambda0, double atweight,
                        int nt, int nrho, double peak_logt, double sigma);
/* Corona for Thomson scattering: the points of a Cartesian m^3 grid over [-3,3] Rsun (in Mm) with
 * 1 < r/Rsun <= 3 (r <= Rsun has no finite scattering factor). n = 1e8 exp(-(r/Rsun-1)/0.07)(1+delta),
 * delta a seeded random-phase field with power ~ k^(-11/3) and rms sigma. With x==NULL only the
 * point count is stored in *count; otherwise the arrays must hold *count entries. */
int fomo_gen_corona(int m, uint64_t seed, double sigma, int64_t* count, float* x, float* y, float* z,
                    float* rho);

/* Delaunay triangulation of n points (host-only; the structure of the tetra render path, which
 * replaces the CGAL triangulation of fomo-CGAL.cpp:33-73). Writes up to cap finite tets as 4
 * input point indices each (positively oriented: det[b-a; c-a; d-a] > 0) and stores the total
 * count in *ntets. Duplicate points appear once, as their last input index. */
int fomo_delaunay(int64_t n, const float* x, const float* y, const float* z, int64_t cap,
                  int64_t* ntets, int32_t* tets);

/* Worker threads of the host Delaunay builder (fomo_delaunay and the tetra path's triangulation;
 * the reference builds its CGAL triangulation serially, fomo-CGAL.cpp:52-57): 0 (the default)
 * takes FOMO_DELAUNAY_THREADS, else OMP_NUM_THREADS, else the hardware threads, at most 64. The
 * tets and their order are the same for every thread count (a canonical order by vertex). */
int fomo_set_delaunay_threads(int threads);
int fomo_delaunay_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* FOMO_HIP_H */
