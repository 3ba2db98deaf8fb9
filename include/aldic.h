/*
 * aldic.h -- C ABI of the MI355X (gfx950) AL-DIC correlation library.
 *
 * Every entry point replaces one function of the reference MATLAB path
 * (jyang526843/2D_ALDIC) and is what a MEX gateway for that function binds; the
 * reference file:line each one stands in for is given beside it.
 *
 * Conventions
 *   - Images are row-major [y][x] (x fastest): the same memory as the reference's
 *     column-major image after its load-time transpose (dim 1 = x, ReadImage.m:79),
 *     so a MATLAB array passes through unchanged.  Coordinates are 1-based.
 *   - Nodal vectors are interleaved like the reference: U = [u1 v1 u2 v2 ...] (2S),
 *     F = [F11 F21 F12 F22 ...] (4S), F = grad u.  H21 holds the lower triangle of
 *     each node's 6x6 Hessian in MATLAB column-major order (S x 21, funICGN.m:320).
 *   - Per-subset status (the reference's ConvItPerEle): 1..100 = converged after
 *     that many residual evaluations, 101 = not converged / left the image / NaN,
 *     102 = more than half of the subset masked, -1 = indexing error (window or
 *     its gradient outside the image), -2 = the mask rule (funICGN.m:57-65) enlarged
 *     the window to round(sqrt(2) ws) > 66 px, beyond the largest subset this library
 *     holds (67 x 67 px), so it was not processed (the reference would process it;
 *     reachable only for ws >= 48 with 10-50 % masked gradients).  Failed subsets
 *     return NaN like the reference's try/catch; the bad-point rules treat -2 as failed.
 *   - Buffers are caller-owned host memory; the library owns device memory.  Calls
 *     are synchronous on return.  One context per host thread.
 *   - Return value: 0 on success, a negative ALDIC_ERR_* code otherwise
 *     (aldic_last_error() gives the message).  No C++ exception crosses the ABI.
 */
#ifndef ALDIC_H
#define ALDIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALDIC_OK 0
#define ALDIC_ERR_ARG -1   /* invalid argument (sizes, NULL, unsupported window) */
#define ALDIC_ERR_HIP -2   /* HIP runtime error */
#define ALDIC_ERR_STATE -3 /* call order (e.g. no images or mesh set) */
#define ALDIC_ERR_NUMERIC -4 /* an iterative solve did not converge within its limit */

enum aldic_dtype { ALDIC_U8 = 0, ALDIC_U16 = 1, ALDIC_F32 = 2, ALDIC_F64 = 3 };

typedef struct aldic_ctx aldic_ctx;

/* Context on HIP device `device`. */
int aldic_create(aldic_ctx **ctx, int device);
void aldic_destroy(aldic_ctx *ctx);
const char *aldic_last_error(const aldic_ctx *ctx);

/* funNormalizeImg.m:22-40 + funImgGradient.m:53-78 (called main_ALDIC.m Section 2).
 * Uploads the raw reference/deformed pair, normalises each by its own ROI mean and
 * N-1 standard deviation (roi = {x1, x2, y1, y2}, 1-based inclusive, clipped;
 * NULL = whole image) and computes the 7-tap gradients of the reference, valid for
 * 4 <= x <= nx-3, 4 <= y <= ny-3, times mask (nx*ny bytes, NULL = all ones). */
int aldic_set_images(aldic_ctx *ctx, const void *f, const void *g, int nx, int ny, int dtype,
                     const uint8_t *mask, const int *roi);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * pre-pass of aldic_set_images (same kernels, same results) on caller-owned DEVICE
 * buffers: raw pair f, g (nx*ny of dtype), mask (nx*ny bytes, NULL = all ones), roi as in
 * aldic_set_images (host int[4], NULL = whole image) -> FP32 planes fo, go (normalised),
 * gxo, gyo (gradients, 0 outside the valid band), nx*ny floats each.  Calls no hipMalloc /
 * hipFree: its work memory (the per-block ROI partial sums and the four statistics,
 * 32,800 bytes) is a fixed device array of the library, one per GPU, looked up by
 * aldic_create; images are limited to 2^29 pixels and 2^22 columns as everywhere (larger
 * ones are rejected with ALDIC_ERR_ARG).  Enqueued on `stream` (a hipStream_t; NULL = the
 * context's stream) and not waited for: synchronise that stream before reading the
 * outputs.  Every call on one GPU shares that GPU's work array, so calls that overlap on
 * one GPU (from several contexts or streams) race. */
int aldic_prep_noalloc(aldic_ctx *ctx, const void *f, const void *g, int nx, int ny, int dtype,
                       const uint8_t *mask, const int *roi, float *fo, float *go, float *gxo,
                       float *gyo, void *stream);

/* The same with the pair already normalised by the caller (the reference's
 * fNormalized/gNormalized, FP64).  Gradients are computed on the device unless
 * aldic_set_gradients() supplies them afterwards. */
int aldic_set_normalized(aldic_ctx *ctx, const double *f, const double *g, int nx, int ny);

/* The reference's Df.DfDx / Df.DfDy (funImgGradient.m:72-78) as full-size nx*ny
 * arrays: gx(x, y) == DfDx(x-3, y-3) inside the valid band, ignored outside. */
int aldic_set_gradients(aldic_ctx *ctx, const double *gx, const double *gy);

/* Device images back to the host (any pointer may be NULL), full size nx*ny. */
int aldic_get_images(aldic_ctx *ctx, double *f, double *g, double *gx, double *gy);

/* funIntegerSearch.m:14-25,90-91: trimmed start corner and subset counts for a
 * whole-field search.  grid = {x1, x2, y1, y2}, s = {sx, sy}.
 * out = {x1', y1', Nx, Ny}.  Host-only. */
int aldic_search_grid(int nx, int ny, const int *grid, int ws, int step, const int *s, int *out);

/* funIntegerSearchWholeField (funIntegerSearch.m:82-200, called by IntegerSearch.m:54
 * with method 1) over the subsets aldic_search_grid() gives.  Outputs are Nx*Ny
 * arrays, entry l*Nx + k = subset (k, l) (x fastest): centres x, y; displacement u, v
 * (findpeak.m subpixel fit, 1e-3 px rounding); cc.max; status (0 ok, -1 constant
 * template -> NaN).  8- and 16-bit images staged with aldic_set_images are searched
 * exactly (int8 matrix cores, integer sums, FP64 NCC); float images in FP32.  qf
 * (Nx*Ny x 2: cc.qfactors, PCE and PPE of the full normxcorr2 map after the per-column
 * normalisation of :174-177; a constant template gives Inf before it) may be NULL to
 * skip the full-map quality factors; the full map must fit 150 KiB of LDS
 * (4 (ws+1)^2 + 4 D^2 + 8 (ws + D)^2 bytes, D = ws + 1 + 2 s). */
int aldic_integer_search(aldic_ctx *ctx, const int *grid, int ws, int step, const int *s,
                         double *x, double *y, double *u, double *v, double *ccmax, double *qf,
                         int *status);

/* funSmoothStrainQuadtree.m:27-104 (called by main_ALDIC_Quadtree.m:172,221,283 with
 * StrainSmoothness 1e-4): two passes of regularizeNd (regularizeNd.m, 'linear', 'normal')
 * of each F component over the grid min:h:max (h = DICmesh.elementMinSize), the
 * nanconv(., fspecial('gaussian', size, std), 'edge', 'nanout') blur and the read-back at
 * the nodes.  nn nodes (nn x 2 coordinates, every one on that grid), F (4 nn, the
 * reference's interleaving) smoothed in place.  filter_std 0 -> 0.5 and filter_size 0 ->
 * 2 ceil(2 std) + 1 as the reference's defaults; only the 3 x 3 filter is built.  The
 * normal equations are solved by FP64 CG (fast-diagonalisation preconditioner) to a
 * relative residual of 1e-12; iters (nullable, 2) = its iterations in each pass. */
int aldic_smooth_strain_quadtree(aldic_ctx *ctx, int nn, const double *coords, double h, double smoothness,
                                 double filter_std, int filter_size, double *F, int *iters);

/* funIntegerSearchMg.m:3-285 (IntegerSearch.m:120, the multigrid search) on the staged
 * images, grid = {x1, x2, y1, y2} (gridxROIRange, gridyROIRange), winsize ws, winstepsize
 * step.  Level 1 correlates the whole ROI (normxcorr2 in full, FFT); each later level
 * splits every region in two or four and correlates it with its g window (+-5 px) on the
 * 13 x 14 lags findpeak reads; the last level's region centres are fitted onto the grid
 * by gridfit ('springs').  dims (int[6]) = {Nx, Ny, levels, final regions, final regions
 * with NaN (the reference's catch), CG iterations of the u/v fit}.  x, y, u, v, ccmax
 * receive Nx*Ny values (x fastest; cap their capacity: ALDIC_ERR_ARG with dims set if
 * smaller, so ((x2-x1)/step+1)*((y2-y1)/step+1) always suffices).  temp (nullable,
 * temp_cap values) receives the final regions' tempx | tempy | tempu | tempv | tempPhi.
 * cc.qfactors is not computed (only the q-factor outlier test reads it). */
int aldic_integer_search_mg(aldic_ctx *ctx, const int *grid, int ws, int step, long cap, int *dims, double *x,
                            double *y, double *u, double *v, double *ccmax, long temp_cap, double *temp);

/* LocalICGN.m:91-107 -> funICGN.m:1-322: 6-DOF IC-GN of S subsets centred at coords
 * (S x 2) from U0 (2S), subset size ws[j] (NULL -> ws_all), stop tolerance tol.
 * Outputs U (2S), F (4S), H21 (S x 21), iters (S); zncc (S, nullable) is the
 * zero-normalised cross-correlation of the final residual evaluation. */
int aldic_icgn6(aldic_ctx *ctx, int S, const double *coords, const double *U0, const int *ws,
                int ws_all, double tol, double *U, double *F, double *H21, int *iters,
                double *zncc);

/* IC-GN 6-DOF as aldic_icgn6, started from a full warp P0 (6 per node: u_x, v_x, u_y, v_y,
 * u, v) instead of (0, 0, 0, 0, U0): the caller-supplied start of the learned initial guess
 * below.  No counterpart in the reference (funICGN.m:39 always starts from zero gradients);
 * for callers that hold an affine start. */
int aldic_icgn6_start(aldic_ctx *ctx, int S, const double *coords, const double *P0,
                      const int *ws, int ws_all, double tol, double *U, double *F, double *H21,
                      int *iters, double *zncc);

/* Subpb1.m:119-152 -> funICGN_Subpb1.m:1-362: 2-DOF IC-GN with F frozen at Fold and
 * the penalty mu * (P(5:6) - Uold - vdual); H21 from LocalICGN.  udual is unused by
 * the reference and is not taken. */
int aldic_icgn2(aldic_ctx *ctx, int S, const double *coords, const int *ws, int ws_all,
                const double *H21, double mu, const double *vdual, const double *Uold,
                const double *Fold, double tol, double *U, int *iters);

/* DICmesh of MeshSetUp.m:80-89: nn nodes (nn x 2 coordinates), E Q4 elements
 * (E x 4, 0-based node indices, counter-clockwise).  Kept on the device for
 * aldic_subpb2 / aldic_nodal_strain. */
int aldic_set_mesh(aldic_ctx *ctx, const double *coords, int nn, const int *elems, int E);

/* Subpb2.m:1-197 (called main_ALDIC.m:287,304,393): assembles
 * (beta K + mu M + alpha mu K) Uhat = beta G (F - udual) + mu M (U - vdual) with
 * 3x3 Gauss points and solves it: directly by fast diagonalisation (FP64 GEMMs with
 * the 1-D generalised eigenbases) on a structured grid (nodes on a uniform lattice, x
 * fastest, the elements of MeshSetUp), else with FP64 Jacobi-PCG to relative residual
 * rtol (maxit iterations at most).  stats (nullable) = {iterations (0 = direct),
 * relative residual}. */
int aldic_subpb2(aldic_ctx *ctx, double beta, double mu, double alpha, const double *U,
                 const double *F, const double *udual, const double *vdual, double rtol,
                 int maxit, double *Uhat, double *stats);

/* The quadtree mesh of main_ALDIC_Quadtree.m (coordinatesFEM, elementsFEM, dirichlet of
 * GenerateQuadtreeMesh.m): nn nodes (nn x 2), E elements (E x 8, 0-based: four corners
 * counter-clockwise, then the midside nodes on the right, top, left and bottom edges,
 * -1 where an edge has none), nd Dirichlet nodes (0-based).  Kept on the device for
 * aldic_subpb2_quadtree. */
int aldic_set_quadtree(aldic_ctx *ctx, const double *coords, int nn, const int *elems, int E,
                       const int *dirichlet, int nd);

/* Subpb2Quadtree.m:1-346 as main_ALDIC_Quadtree.m:181-212 calls it (alpha = 0, 2x2 Gauss
 * points, duals multiplied by 0): solves (beta K + mu M) Uhat = beta G F + mu M U over
 * the free nodes with FP64 Jacobi-PCG to relative residual rtol (maxit iterations at
 * most); Uhat = U at the Dirichlet nodes and 0 at nodes outside every element and at the
 * first (lowest) element node when every element has all four midside nodes (the
 * reference's free-node list assumes a 0 entry there).  U: 2 nn, F: 4 nn
 * ([F11 F21 F12 F22] per node); stats (nullable) = {iterations, relative residual}. */
int aldic_subpb2_quadtree(aldic_ctx *ctx, double beta, double mu, const double *U,
                          const double *F, double rtol, int maxit, double *Uhat, double *stats);

/* funGlobalNodalStrainQuadtree.m:31-255 (GaussPtOrder 2, smoothness 0) on the quadtree
 * mesh of aldic_set_quadtree: the midside nodes are dropped (:33); per element the 2x2
 * Gauss-point strains [u_x u_y v_x v_y] (:184-197, the reference's slot order) at
 * x_i = g_i (x3 - x1)/2 + (x3 + x1)/2, y_j likewise, row 4 e + 2 i + j (:116-121,
 * :201-205) -> Fgp (4E x 4) and gpxy (4E x 2), both nullable.  The reference then
 * evaluates scatteredInterpolant(gpxy, Fgp(:,k), 'linear') at the nodes (:246-255); here
 * the caller passes that interpolation as weights: node i gets
 * sum_k w[3i+k] Fgp[tri[3i+k]] -> F (4 nn, [F11 F21 F12 F22] slots holding
 * [u_x u_y v_x v_y] as the reference stores them).  tri / w / F nullable together. */
int aldic_nodal_strain_quadtree(aldic_ctx *ctx, const double *U, const int *tri, const double *w,
                                double *F, double *Fgp, double *gpxy);

/* funGlobal_NodalStrainAvg.m:1-147: nodal gradient by 2x2 Gauss points,
 * extrapolation and element averaging.  swap_fix = 0 writes the reference's slot
 * order [u_x u_y v_x v_y] (off-diagonals swapped relative to F's convention);
 * swap_fix = 1 writes [u_x v_x u_y v_y]. */
int aldic_nodal_strain(aldic_ctx *ctx, const double *U, int swap_fix, double *F);

/* ComputeStrain.m:11-139 (main_ALDIC.m:611 runs it per frame) on a uniform M x N node
 * grid, node n = i + M j (x fastest, MeshSetUp's order), U interleaved (2 M N), F per
 * node (F11, F21, F12, F22).  method 0: the solved F (ComputeStrain.m:12-24);
 * 1: funDerivativeOp.m's finite differences of U with step h (central inside,
 * one-sided on the border; :27-42); 2: those, then PlaneFit2 plane fits over
 * (2 rad + 1)^2 nodes for the nodes at least rad from the border (:45-76; the
 * reference asks for rad interactively); 3: funGlobal_NodalStrainAvg of U on the mesh
 * set by aldic_set_mesh, which must be this grid (:79-94, its slot order kept).
 * strain_type 0 infinitesimal, 1 Eulerian, 2 Green-Lagrangian (:98-136; 3, "code by
 * yourself", is rejected).  Ffinite (4 M N) receives FStrainFinite; Fworld (nullable,
 * 4 (M - 2r)(N - 2r)) the trimmed FStrainWorld (nodes at least r from the border, r = 0,
 * 1, rad, 1 for methods 0..3, x fastest, F21 and F12 negated).  Replaces the
 * ComputeStrain.m script (no reference function signature). */
/* inpaint_nans.m method 3 (del^4 plate) or 4 (springs) of a rows x cols array in
 * MATLAB's column-major layout, in place (as Init.m:13 and funRemoveOutliers.m:58,375
 * call it): the NaNs get the least-squares solution of the method's stencil rows with
 * the known values fixed, solved on the device by a banded Cholesky of the normal
 * equations.  ALDIC_ERR_NUMERIC if every element is NaN or the band would exceed
 * 512 MiB. */
int aldic_inpaint_nans(aldic_ctx *ctx, double *A, int rows, int cols, int method);

/* funRemoveOutliers(u, v, cc, qDICOrNot != 1, Thr0) (funRemoveOutliers.m:45-86) without
 * its prompts, on two rows x cols column-major arrays in place: the single-precision
 * normalised median test (funMedRemoveOutliers, medFilt2 [2 2] 'symmetric' skipping the
 * second neighbour, NaNs pre-filled by method 3), elements with fluctuation magnitude
 * above thr set NaN and in-painted by springs, the original NaNs put back.  flagged
 * (nullable, rows x cols column-major) receives RemoveOutliersList as 0/1; n_flagged
 * (nullable) its count.  The q-factor removal (qDICOrNot == 1) and the interactive bound
 * and click removal are not provided. */
int aldic_remove_outliers(aldic_ctx *ctx, double *u, double *v, int rows, int cols, double thr,
                          int *flagged, int *n_flagged);

int aldic_compute_strain(aldic_ctx *ctx, int M, int N, double h, const double *U,
                         const double *F, int method, int rad, int strain_type, double *Ffinite,
                         double *Fworld);

/* LocalICGN.m:141-161, Subpb1Quadtree.m:188-199: scatteredInterpolant(good nodes, values,
 * 'natural') evaluated at every node.  nn nodes (nn x 2 coordinates), bad[i] != 0 marks a
 * node with no datum; U (2 nn) and F (4 nn, nullable) are refilled in place at the bad
 * nodes (good nodes keep their values).  Inside the good nodes' convex hull: Sibson's
 * natural-neighbour interpolant (Voronoi area stolen by the node); on or outside it (where
 * MATLAB extrapolates 'linear'ly from "boundary gradients", which is not documented
 * further): the plane through the good nodes next to the nearest hull point and the
 * Delaunay third vertex of that hull edge (at a hull vertex: the vertex and its two hull
 * neighbours); fewer than three non-collinear good nodes: the nearest good node; none:
 * NaN.  ALDIC_ERR_NUMERIC if a Voronoi cell could not be formed (its node gets NaN). */
int aldic_natural_fill(aldic_ctx *ctx, int nn, const double *coords, const int *bad, double *U,
                       double *F);

/* ba_interp2.cpp:318-335 (funICGN.m:125): Keys-cubic samples of the deformed image
 * at (X[i], Y[i]) (X along dim 1), clamp-to-edge. */
int aldic_bicubic(aldic_ctx *ctx, const double *X, const double *Y, long n, double *Z);

/* gridfit.m as Subpb1 calls it (Subpb1.m:185-189: 'regularizer','springs', other
 * options default) for data on the nodes of the gx x gy grid xnodes x ynodes (both
 * strictly increasing): z[iy*gx + ix] is the datum at node (ix, iy), NaN where there
 * is none.  out[iy*gx + ix] receives zgrid(iy, ix) (gridfit.m:279, the least-squares fit
 * of gridfit.m:651-666).  FP64 conjugate gradients to rtol (relative residual of the
 * normal equations, e.g. 1e-12) in at most maxit iterations; stats (nullable) =
 * {iterations, relative residual}. */
int aldic_gridfit(aldic_ctx *ctx, int gx, int gy, const double *xnodes, const double *ynodes,
                  const double *z, double rtol, int maxit, double *out, double *stats);

/* ---- one whole frame on the device (main_ALDIC.m:87-451, FEM branch) ----------
 * Raw images are staged once; aldic_frame_run() then runs, without host
 * round-trips: normalisation + gradients, integer search, 6-DOF IC-GN with the
 * LocalICGN bad-point rule (LocalICGN.m:119-139) and refill, the step-1 Subpb2 +
 * nodal strain + dual update at fixed beta, and n_admm iterations of
 * (2-DOF IC-GN with Subpb1's bad-point rule and gridfit refit of every node
 * (Subpb1.m:165-199) -> Subpb2 -> nodal strain -> dual update)
 * (main_ALDIC.m:357-451).  Smoothing (funSmoothDisp/funSmoothStrain) runs only with
 * the `smooth` option below.
 * A gridfit runs a budget of CG iterations (the most the previous frame needed plus
 * a margin) without a host wait; if one needs more, the frame runs a second time with
 * gridfits that poll for convergence, so the result never depends on the budget. */
typedef struct {
    int ws, step;
    int s[2];          /* search radius */
    int grid[4];       /* x1 x2 y1 y2 ROI for the search grid */
    double tol;        /* IC-GN stop tolerance (1e-2) */
    double mu, beta;   /* ADMM penalties */
    int n_admm;        /* ADMM iterations after step 1 */
    int cg_maxit;      /* PCG iterations per Subpb2 */
    double cg_rtol;    /* PCG relative residual */
    int no_subpb1_fit; /* 1: skip Subpb1's gridfit refit (bad nodes are only refilled) */
    int stage_times;   /* 1: an event at every stage boundary (all of aldic_frame_times);
                        * 0: events only around the search, the 6-DOF IC-GN and the
                        * whole frame (each event costs the stream a few microseconds) */
    int smooth;        /* 1: main_ALDIC.m's smoothing passes (funSmoothStrain.m, funSmoothDisp.m:
                        * two passes of gridfit 'springs' over the node grid + 3x3 Gaussian
                        * sigma 0.5 with edge correction): F1 after LocalICGN (:233), F2 after
                        * every global step (:320,:407), U1 and U2 in the ADMM steps (:381,:400) */
    int beta_sweep;    /* 1: beta from the step-1 global step at the six candidates
                        * [1e-3, 1e-5^.5, 1e-2, 1e-3^.5, 1e-1, 1e-1^.5] step^2 mu (main_ALDIC.m:
                        * 237,274-301, FE branch); the beta field is then unused */
    double admm_tol;   /* > 0: stop the ADMM steps once the RMS change of U2 or of Subpb1's U
                        * (before smoothing) between steps is below it (main_ALDIC.m:409-447,
                        * tol2 = 1e-4); n_admm stays the most steps (the reference: 4) */
    double outlier_thr; /* > 0: after the integer search, nodes whose centre pixel is masked
                        * become NaN (IntegerSearch.m:68-76), the normalised median test of
                        * funRemoveOutliers at Thr0 = outlier_thr (the reference passes 100,
                        * IntegerSearch.m:114) flags outliers, and the flagged and NaN nodes
                        * are in-painted by springs (inpaint_nans method 4; the same solve as
                        * funRemoveOutliers' followed by Init.m:13's) before LocalICGN;
                        * 0: the search's field goes to LocalICGN as it is */
    int learned_start; /* 1: the learned initial guess (aldic_cnn_*, weights set) starts
                        * LocalICGN's 6-DOF IC-GN from its full affine warp in place of the
                        * integer search; subsets whose IC-GN misses status 1..20 or ends
                        * with ZNCC below 0.9 are searched (FP32 search on that list) and
                        * run again from the search's peak, so converged IC-GN still decides
                        * every displacement; the search's outlier stage does not run.
                        * Not for aldic_frame_run_seeded, which starts from the caller's U. */
} aldic_frame_opts;

int aldic_frame_stage(aldic_ctx *ctx, const void *f, const void *g, int nx, int ny, int dtype,
                      const uint8_t *mask, const int *roi);
int aldic_frame_setup(aldic_ctx *ctx, const aldic_frame_opts *opts, int *n_nodes);
int aldic_frame_run(aldic_ctx *ctx);
/* main_ALDIC.m:134-137 (a later frame of a sequence with DICpara.NewFFTSearch = 0, images
 * 3..6): the same frame with the integer search replaced by the caller's initial guess
 * U0 (2S, interleaved; e.g. the previous frame's U, ResultDisp{ImgSeqNum-2}.U). */
int aldic_frame_run_seeded(aldic_ctx *ctx, const double *U0);
/* Results of the last frame: U, F (after the last global step), U6/F6/iters6 of the
 * 6-DOF IC-GN before post-processing, iters of the last 2-DOF step; any may be NULL. */
int aldic_frame_get(aldic_ctx *ctx, double *U, double *F, double *U6, double *F6, int *iters6,
                    int *iters2, double *u0);
/* Stage times of the last frame in ms: {prep, search, icgn6, post6, subpb2 total,
 * icgn2 total, strain+dual total, frame, search kernel, 6-DOF IC-GN kernel}; with
 * stage_times 0 only search, icgn6, frame and the two kernels are measured (the others
 * read 0).  The search kernel's time excludes the window box-sum planes built before it;
 * the 6-DOF kernel's excludes the re-launch of windows enlarged by the mask rule.
 * n <= 10. */
int aldic_frame_times(aldic_ctx *ctx, double *ms, int n);
/* Counters of the frame driver (no counterpart in the reference): {iteration kernels
 * each Subpb1 gridfit of the next frame runs before its convergence is checked,
 * frames run a second time because a gridfit needed more, largest gridfit iteration
 * count of the last frame}, then for ADMM steps 1..8 the bad-node count of that step's
 * Subpb1 (statuses < 0 or > 99, Subpb1.m:166-169; -1 for a step not run), which with
 * LocalICGN's count (from iters6, LocalICGN.m:117-122) is ALSub1BadPtNum
 * (main_ALDIC.m:180,365,456). n <= 11. */
int aldic_frame_counters(aldic_ctx *ctx, int *out, int n);
/* (out[11], n = 12: subsets of the last frame that fell back from the learned start to
 * the search, -1 when the frame did not use the learned start.) */
/* Scalars of the last frame (no counterpart in the reference's outputs; main_ALDIC.m
 * prints them): {beta used, ADMM steps run, RMS change of U2 in the last step, RMS change
 * of Subpb1's U in the last step (both measured only with admm_tol > 0, else NaN)}.
 * n <= 4. */
int aldic_frame_scalars(aldic_ctx *ctx, double *out, int n);

/* Quadtree frame: main_ALDIC_Quadtree.m:159-312 (Sections 4-5 of one frame, smoothing
 * off) on the mesh of aldic_set_quadtree (its nodes are the subset centres, its
 * Dirichlet nodes the hole-edge nodes) and the images of aldic_set_images:
 *   LocalICGN (6-DOF, window ws[i] per node; LocalICGN.m:119-131 bad-point rule),
 *   Subpb2Quadtree at beta with zero duals, funGlobalNodalStrainQuadtree, hole-edge F
 *   kept from Subpb1 (:222,:225), udual = F2 - F1, vdual = U2 - U1, then n_admm times
 *   Subpb1Quadtree (2-DOF on the same per-node windows, its winsizeList of
 *   Subpb1Quadtree.m:50, with LocalICGN's Hessian; the max(mean + 0.15 std, 20) rule;
 *   F1 = F2), Subpb2Quadtree, nodal strain, hole-edge F, duals (:251-312).
 * Bad nodes are refilled by natural-neighbour (Sibson) interpolation of the good nodes,
 * as the reference's scatteredInterpolant 'natural' (see aldic_natural_fill).  The nodal strain uses
 * the Gauss-point interpolation weights tri / w (nn x 3) as in
 * aldic_nodal_strain_quadtree. */
typedef struct {
    double tol;        /* IC-GN stop tolerance (1e-2) */
    double mu, beta;   /* ADMM penalties */
    int n_admm;        /* ADMM steps after step 1 (3) */
    int cg_maxit;      /* PCG iterations per global step */
    double cg_rtol;    /* PCG relative residual */
    double smooth_h;   /* > 0: funSmoothStrainQuadtree as main_ALDIC_Quadtree.m:170-172,
                        * 221, 283-284 call it (on F1 after LocalICGN, on 0.1 F2 + 0.9 F1
                        * after every global step, hole-edge F copied before and, in the
                        * ADMM loop, after), with elementMinSize smooth_h (every node on
                        * the grid min:h:max) and the 3 x 3 default filter; 0: off */
    double smoothness; /* StrainSmoothness (the reference sets 1e-4) */
} aldic_qframe_opts;
int aldic_qframe_setup(aldic_ctx *ctx, const aldic_qframe_opts *opts, const int *ws,
                       const int *tri, const double *w);
/* One frame from the initial guess U0 (2 nn). */
int aldic_qframe_run(aldic_ctx *ctx, const double *U0);
/* Results of the last frame, any may be NULL: U (2 nn) and F (4 nn) after the last
 * global step, iters6 of LocalICGN, iters2 of the last Subpb1, times = {frame ms,
 * LocalICGN + bad-point step ms, global steps ms (Subpb2 + strain + duals, all steps;
 * NaN unless the frame ran with polled PCGs), PCG iterations (all steps), PCG iteration
 * kernels enqueued per global step without a host wait (sized from the previous frame),
 * frames re-run with polled PCGs because a step needed more}.  The frame waits on the
 * host once, at its end; a PCG that did not converge in its budget re-runs the frame. */
int aldic_qframe_get(aldic_ctx *ctx, double *U, double *F, int *iters6, int *iters2,
                     double *times);

/* ---- learned initial guess (no counterpart in the reference) ------------------------
 * A siamese correlation network on the BF16 matrix cores proposes each subset's six IC-GN
 * parameters (u_x, v_x, u_y, v_y, u, v) in place of the integer search: per node the
 * 64 x 64 patches of the normalised f and g centred on it, plus a 4-pixel context ring
 * (72 x 72); a shared encoder on 4 x 4 pixel cells (patchify to 16 channels + ReLU, 3 x 3
 * 'valid' convolution), features normalised per cell, cost volumes of the four quadrants
 * of the central 24 x 24 pixels over displacements of -5..5 cells, a soft-argmax of each
 * (and of their mean) and a dense 10 -> 6 head.  Weights (FP32, aldic_cnn_nparams() of
 * them) are set by the caller or trained here on a device-held set of patch pairs with
 * targets (mean squared error on (F / 0.05, U / 4), Adam).  Pixels outside the image are
 * clamped.  On node sets whose inputs share one 4-pixel lattice (grids with a step that
 * is a multiple of 4) the encoder runs once per image. */
long aldic_cnn_nparams(void);
int aldic_cnn_set_weights(aldic_ctx *ctx, const float *w, long n);   /* resets Adam state */
int aldic_cnn_get_weights(aldic_ctx *ctx, float *w, long n);
/* P (6 per node) at nodes coords (1-based centres, S x 2) of the current images. */
int aldic_cnn_predict(aldic_ctx *ctx, int S, const double *coords, double *P);
/* Appends the nodes' patch pairs of the current images, with targets P (6 per node), to
 * the training set; aldic_cnn_clear_samples empties it. */
int aldic_cnn_add_samples(aldic_ctx *ctx, int S, const double *coords, const double *P);
int aldic_cnn_clear_samples(aldic_ctx *ctx);
/* steps Adam steps of batch rows drawn (with replacement, splitmix64 from seed) from the
 * training set; loss[i] = the mean squared error of step i before its update. */
int aldic_cnn_train(aldic_ctx *ctx, int steps, int batch, float lr, unsigned long long seed,
                    double *loss);
/* Loss, gradient (nparams) and scaled outputs (6 B, nullable) of training-set rows, no
 * update (for gradient checks). */
int aldic_cnn_grad(aldic_ctx *ctx, int B, const int *rows, float *grad, double *loss,
                   double *outputs);
/* LocalICGN's start and 6-DOF IC-GN (LocalICGN.m:91-117, without the bad-point refill)
 * over the whole-field search grid of aldic_integer_search (grid, ws, step, s as there),
 * resident on the device.  start 0: from the integer search (IntegerSearch.m:54).  start
 * 1: from the learned warp of every node, then, on the subsets whose IC-GN did not reach
 * status 1..max_it or ended with ZNCC below min_zncc (fallback[i] = 1), the FP32 integer
 * search and IC-GN from its peak.  Outputs per node (S = Nx * Ny, x fastest): x, y
 * (nullable), U (2S), F (4S), iters, zncc and fallback (nullable); n_fallback (nullable);
 * times_ms (5, nullable): start (search or network), IC-GN, fallback search, fallback
 * IC-GN, total, from HIP events on the library's stream. */
int aldic_local_icgn(aldic_ctx *ctx, int start, const int *grid, int ws, int step, const int *s, double tol,
                     int max_it, double min_zncc, double *x, double *y, double *U, double *F, int *iters,
                     double *zncc, int *fallback, int *n_fallback, float *times_ms);

#ifdef __cplusplus
}
#endif
#endif
