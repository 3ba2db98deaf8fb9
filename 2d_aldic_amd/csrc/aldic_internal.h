// Internal declarations shared by the gfx950 kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/aldic.h"

namespace aldic {

// One (node, element) incidence: the element, the node's local index in it and the
// CSR slots of (node, element node b) for b = 0..3.
struct N2E {
    int e, a;
    int slot[4];
};

// The same for the quadtree's 8-node elements (corners 0..3, midside nodes 4..7 that
// may be absent); slot[b] is unused where element node b is absent.
struct N2E8 {
    int e, a;
    int slot[8];
};

// Device buffer that only grows.
struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
        if (e == hipSuccess) cap = bytes;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Node set of a natural-neighbour refill: buckets of side h (CSR by bucket, x fastest)
// and rows of equal y (x ascending), plus the refill's work memory.
struct NatGrid {
    int nn = 0, BX = 0, BY = 0, R = 0;
    double ox = 0, oy = 0, h = 1;
    DBuf bstart, bnode, rstart, rnode, work, big;  // big: second-pass Sibson polygons
};
hipError_t natgrid_setup(NatGrid &g, const double *coords_host, int nn, hipStream_t st);
// LocalICGN.m:141-161 / Subpb1Quadtree.m:188-199 refill in place: bad[i] != 0 nodes of U
// (2 per node) and F (4 per node, nullable) from the good ones (scatteredInterpolant
// 'natural', linear rule on and outside the good nodes' hull).  No host wait.
// err (nullable, device): refills whose cell could not be formed are added to *err.
hipError_t natural_fill(const NatGrid &g, const double *coords_dev, const int *bad, double *U,
                        double *F, hipStream_t st, int *err = nullptr);
hipError_t natural_err_clear(int *err, hipStream_t st);
// Refills of the last natural_fill on g whose Sibson cell could not be formed (waits).
int natural_errors(const NatGrid &g, hipStream_t st);

// Q4 mesh in node-row (CSR-like) form: for node i, the elements around it and the
// node's local index in each; plus the scalar node-to-node sparsity pattern (the
// u- and v-blocks of the global matrix share it).
struct MeshDev {
    int nn = 0, E = 0;
    int nnz = 0;              // scalar pattern entries
    DBuf coords;              // double2[nn]
    DBuf elems;               // int4[E]
    DBuf n2e_ptr, n2e;        // int[nn+1], int2[...] = (element, local index)
    DBuf rowptr, col;         // scalar CSR pattern, int[nn+1], int[nnz]
    DBuf aval, mval, diag;    // double[nnz] (beta K + mu M (+alpha mu K)), double[nnz] M, double[nn]
    DBuf gvec;                // double2[nn]: sum over elements of int grad N_i dA
    double beta = -1, mu = -1, alpha = -1;  // values aval was assembled for
    // Structured grid (nodes x-fastest on a uniform gx x gy lattice, standard Q4
    // elements): K and M are Kronecker sums/products of the 1-D matrices, solved
    // directly by fast diagonalisation (fem_solve).
    bool structured = false;
    int gx = 0, gy = 0;
    double hx = 0, hy = 0;
    DBuf fdV;                 // double[gx*gx + gy*gy]: M-orthonormal eigenbases Vx, Vy
    DBuf fdlam;               // double[gx + gy]: generalised eigenvalues of (K1, M1)
};

// Quadtree mesh of the global step: node rows of an 8-node-element mesh in m (coords,
// n2e_ptr, CSR pattern, aval / mval / diag / gvec), the elements (int[8 E], -1 = no
// midside node) and their incidences, each node's role code (0 free, 1 Dirichlet =
// kept at U, 2 fixed at 0) and cval = the matrix entries of free rows in fixed columns.
struct QuadDev {
    MeshDev m;
    DBuf elems, n2e, code, cval;
    NatGrid nat;  // its nodes for the bad-node refill
};

// gridfit ('springs') of nodal data on a rectilinear node grid, data on the nodes
struct GridfitDev {
    int gx = 0, gy = 0;
    double lam2 = 0.0;
    DBuf w;        // squared stiffnesses [x springs | y springs | (dx/xs)^2 | (dy/ys)^2]
    DBuf a9, dinv; // 9-point stencil of N (slot-major), inverse of its diagonal
};
hipError_t gridfit_setup(GridfitDev &g, const double *xnodes, int gx, const double *ynodes,
                         int gy, hipStream_t st);
// U: np (1 or 2) interleaved value pairs per node (node = iy * gx + ix), the initial
// guess on entry and the fit of all on return (one CG on them together); bad[node] != 0:
// the node carries no data (bad may be null: every node does).  stats_dev (nullable, device): {converged (1) or stopped on a breakdown (1)
// or not (0), iterations, relative residual of the pair}.  work[0..1] are used.
// poll > 0: a host check for convergence every poll iterations, up to maxit; poll <= 0:
// exactly maxit iteration kernels (those after convergence return at once) and no host
// wait, so the caller reads stats_dev to learn whether maxit sufficed.
hipError_t gridfit_solve(GridfitDev &g, const int *bad, double *U, int np, double rtol, int maxit,
                         DBuf *work, double *stats_dev, hipStream_t st, int poll);
// gridfit.m ('springs', 'triangle', smoothness 1) of np scattered points (px, py device,
// strictly inside the grid; pd their npair (1 or 2) value pairs each) onto the grid g was
// set up for (xnodes_dev, ynodes_dev: its nodes on the device): U (gx * gy nodes of npair
// pairs, node iy * gx + ix) receives the fit of every pair (one CG on them together).
// work[0..2] are used; a9 is overwritten.
hipError_t gridfit_scattered(GridfitDev &g, const double *xnodes_dev, const double *ynodes_dev,
                             int np, const double *px, const double *py, const double2 *pd, int npair,
                             double *U, double rtol, int maxit, DBuf *work, double *stats_dev,
                             hipStream_t st);
// F2 = 0.1 F2 + 0.9 F1 (main_ALDIC_Quadtree.m:221,283), n values
hipError_t launch_blend_f(double *F2, const double *F1, long n, hipStream_t st);
// the grid min:h:max of node coordinates (host, nn x 2): nx, ny and each node's grid index
// iy * nx + ix; false if a node is off the grid or the grid is too small or too large
bool smooth_grid(int nn, const double *coords, double h, int &nx, int &ny, std::vector<int> &gidx);
// funSmoothStrainQuadtree's two passes on F (4 per node, device, in place): nn nodes on
// the nx x ny grid at grid indices gidx (device); w[0..1] work; iters[2] CG iterations
hipError_t smooth_strain_qt(int nn, const int *gidx_dev, int nx, int ny, double smoothness, double sigma, double rtol,
                            int maxit, double *F, DBuf *w, int *iters, hipStream_t st);
// nanconv(., fspecial('gaussian', 3, sigma), 'edge') of nc values per node on the gx x gy
// node grid (out-of-place)
hipError_t launch_blur3(const double *in, double *out, int gx, int gy, int nc, double sigma,
                        hipStream_t st);

// Quadtree frame (main_ALDIC_Quadtree.m's Sections 4-5 on the device).
struct QFrameState {
    bool ready = false;
    aldic_qframe_opts o{};
    int S = 0;
    DBuf ws, U0, U6, F6, H21, it6, U1, F1, U2, F2a, F2b, udual, vdual, it2, bad, part, known,
        tri, w, Fgp, lists, rl, b, cc;
    std::vector<int> boff, bn, bppl;  // 6-DOF launches: list offset, length and pixel class
    int pcg_iters = 0;                // PCG iterations of the last frame's global steps
    int cur_F2 = 0;                   // which of F2a / F2b holds the last F
    int pcg_budget = 64;              // PCG iterations enqueued per global step unpolled
    int pcg_reruns = 0;               // frames re-run with polled PCG
    int pcg_most = 0;                 // most PCG iterations of one global step, last frame
    DBuf sgidx, sw[2];                // strain smoothing: node -> grid index, work
    int snx = 0, sny = 0;             // its grid
    int smooth_iters = 0;             // its CG iterations in the last frame
    DBuf pstat;                       // {converged, iterations, residual} per global step
    DBuf nat_err;                     // refills whose cell could not be formed (device int)
    int *nat_err_h = nullptr;         // pinned copy
    double *hstat = nullptr;          // pinned host copy
    double t_frame = 0, t_local = 0, t_global = 0;  // ms of the last frame
    hipEvent_t ev[4] = {};
};

// ADMM steps whose Subpb1 bad-node counts a frame records (ALSub1BadPtNum holds six
// entries, main_ALDIC.m:174; the first is LocalICGN's).
constexpr int kMaxSteps = 8;

// The learned start's launch plan for a node set (cnn.hip): the lattice path when every
// node shares one 4-pixel cell lattice.
struct CnnPlan {
    bool lattice = false;
    int px0 = 0, py0 = 0;
    long nx1 = 0, ny1 = 0;
};

struct FrameState {
    bool ready = false;
    aldic_frame_opts o{};
    int x1 = 0, y1 = 0, Nx = 0, Ny = 0, S = 0;
    DBuf coords, U0, U6, F6, H21, it6, zncc, Ucur, Fcur, U1, F1, U2, F2, udual, vdual, it2,
        bad, tmp, part, cc, rl, F2b, gfstat,
        Us, Fs, U1s, U1old, U2old, errs;  // smoothing temporaries, stop-rule copies, norms
    GridfitDev gf;  // Subpb1's gridfit over the node grid
    NatGrid nat;    // the node grid for LocalICGN's bad-node refill
    DBuf seed;      // aldic_frame_run_seeded: the caller's initial guess (2S)
    DBuf nat_err;   // refills of the frame whose cell could not be formed (device int)
    DBuf nbad;      // bad-node count of each ADMM step's Subpb1 (device int[kMaxSteps], -1: not run)
    int n_outliers = 0;  // nodes the last searched frame's outlier test flagged
    CnnPlan plan;        // learned_start: the network's plan for the node grid
    DBuf P6, F0, lflag;  // learned_start: warps (6S), F0 (4S), fallback flags, list, count
    int n_fallback = -1; // subsets of the last frame that fell back to the search (-1: none run)
    int *nfb_h = nullptr;  // pinned copy of the fallback count
    int *nat_err_h = nullptr;  // pinned copy, read after the frame's wait
    int f2cur = 0;  // which of F2 / F2b holds the last global step's F
    double *gfh = nullptr;  // pinned: {converged, iterations, residual} per frame gridfit
    double *hostd = nullptr;  // pinned: norms read back by the beta sweep and stop rule
    int gf_cap = 0;         // gridfit records a frame can hold
    int gf_count = 0;       // gridfits the last frame pass ran
    double beta_used = 0.0, upd = 0.0, upd2 = 0.0;  // beta, last RMS updates of U2 and U1
    int admm_steps = 0;     // ADMM steps the last frame ran
    int gf_budget = 16;     // iteration kernels per gridfit before a frame is re-run
    int gf_reruns = 0;      // frames re-run with polled gridfits
    int gf_most = 0;        // most gridfit iterations in the last frame
    std::vector<double> hcoords;
    std::vector<int> helems;
    std::vector<hipEvent_t> ev;  // stage-boundary events of frame_run
    hipEvent_t ev_skern = nullptr, ev_ikern = nullptr;  // search kernel start, 6-DOF kernel end
    float ms[10] = {0};
};

struct OutlierWork;  // outlier.hip
struct CnnWork;      // cnn.hip
struct MgWork;       // search_mg.hip

}  // namespace aldic

struct aldic_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int nx = 0, ny = 0;
    bool have_images = false;
    aldic::DBuf f, g, gx, gy;        // FP32 normalised images and gradients
    aldic::DBuf rawf, rawg, mask;    // staged raw input
    int raw_dtype = 0;
    int roi[4] = {0, 0, 0, 0};
    bool have_mask = false;
    bool images_from_raw = false;    // f, g normalised on the device from rawf, rawg
    aldic::OutlierWork *outl = nullptr;  // work buffers of the outlier test and in-painting
    aldic::CnnWork *cnn = nullptr;       // learned initial guess: weights, training set
    aldic::MgWork *mg = nullptr;         // multigrid integer search
    aldic::DBuf stats;               // ROI partial sums, then {mean_f, 1/std_f, mean_g, 1/std_g}
    aldic::DBuf boxp;                // window box-sum planes of the exact search
    aldic::DBuf scratch[9];          // per-call work buffers (8: IC-GN start gradients)
    aldic::DBuf lic[8];              // aldic_local_icgn work buffers (kept between calls)
    aldic::DBuf smq[4];              // aldic_smooth_strain_quadtree work buffers
    aldic::MeshDev mesh;
    bool have_mesh = false;
    aldic::QuadDev quad;             // aldic_set_quadtree / aldic_subpb2_quadtree
    bool have_quad = false;
    aldic::FrameState fr;
    aldic::QFrameState qf;
    aldic::GridfitDev gf;            // aldic_gridfit
    aldic::NatGrid nat;              // aldic_natural_fill
    hipEvent_t ev[10] = {};
};

namespace aldic {

// ---- kernels' host launchers (all enqueue on ctx->stream) ---------------------
constexpr int kStatParts = 1024;
// Largest image: FP32 planes are addressed with 32-bit byte offsets in the kernels,
// and row offsets (4 nx bytes) with 24-bit multiplies, so nx < 2^22.
constexpr size_t kMaxPixels = (size_t)1 << 29;  // per-block partials written by launch_stats
// both images' partials in one launch: acc = 2 kStatParts of f, then 2 kStatParts of g
hipError_t launch_stats(const void *rawf, const void *rawg, int dtype, int nx, int ny,
                        const int *roi, double *acc, hipStream_t st);
hipError_t launch_prep(const void *rawf, const void *rawg, int dtype, int nx, int ny,
                       const uint8_t *mask, const double *acc, double n_roi, double *ms,
                       float *f, float *g, float *gx, float *gy, hipStream_t st);
hipError_t launch_gradient(const float *f, int nx, int ny, float *gx, float *gy, hipStream_t st);
// Looks up the device address of aldic_prep_noalloc's work array on `device` (current).
hipError_t noalloc_init(int device);

struct IcgnArgs {
    const float *f, *g, *gx, *gy;
    int nx, ny, S;
    const double *coords;   // S x 2
    const int *ws;          // S or nullptr
    int ws_all;
    double tol;
    // 6-DOF
    const double *U0;       // 2S
    const double *F0;       // 4S or nullptr: starting displacement gradient (else 0)
    double *U, *F, *H21, *zncc;
    int *iters;
    // 2-DOF
    const double *Hin, *vdual, *Uold, *Fold;
    double mu;
    const int *list;        // subset indices to process (nullptr = all S)
    int nlist;
    const int *nlist_dev;   // when set, the list length is read from the device
    double *Ucopy, *Fcopy;  // 6-DOF: optional second copy of U and F
    // per-subset constants shared by the 6-DOF and 2-DOF runs on the same windows
    // (8 doubles per subset); cc_mode 0: unused, 1: written by 6-DOF, 2: read by 2-DOF
    double *cc;
    int cc_mode;
};
hipError_t launch_icgn6(const IcgnArgs &a, int ppl, hipStream_t st);
hipError_t launch_icgn2(const IcgnArgs &a, int ppl, hipStream_t st);
int icgn_ppl_for(int ws);   // pixels per lane class that holds a (ws+1)^2 window, 0 if none
// list = indices j (ascending) with it[j] == code, *count = their number; one workgroup
// list: int[S], count: int[1], scratch: int[(S + 4095) / 4096]
hipError_t launch_list_eq(const int *it, int S, int code, int *list, int *count, int *scratch,
                          hipStream_t st);

struct SearchArgs {
    const float *f, *g;
    int nx, ny, x1, y1, Nx, Ny, ws, step, sx, sy;
    double *x, *y, *u, *v, *ccmax, *qf;
    int *status;
    double *U0;             // optional interleaved output (2S), may be nullptr
    // raw 8- or 16-bit pair and {mean_f, 1/std_f, mean_g, 1/std_g} of the normalisation
    // that made f, g: when set (and the window fits) the exact int8 matrix-core search
    // runs; planes is its work buffer for the window box sums (large lag maps)
    const uint8_t *f8 = nullptr, *g8 = nullptr;
    const uint16_t *f16 = nullptr, *g16 = nullptr;
    const double *ms = nullptr;
    DBuf *planes = nullptr;
    hipEvent_t ev_kernel = nullptr;  // recorded just before the main search kernel (nullable)
    // FP32 search on a subset list only (launch_search_list): grid indices list[i] for
    // i < *nlist_dev, at most Nx * Ny of them; outputs land at those indices
    const int *list = nullptr, *nlist_dev = nullptr;
};
// Runs an exact int8 matrix-core search when raw pixels are given: k_search_i8 for 8-bit
// images with 2s+3 <= 32 lags per axis and ws + 2s + 3 <= 64, k_search_mt for 16-bit
// images or larger lag maps (ws <= 64); else the FP32 search on a.f, a.g.
hipError_t launch_search(const SearchArgs &a, hipStream_t st);
// The FP32 search (k_search on a.f, a.g) over the first n subsets a.list names (n at
// most *a.nlist_dev).
hipError_t launch_search_list(const SearchArgs &a, int n, hipStream_t st);
size_t search_lds_bytes(int ws, int sx, int sy);
bool search_i8_supported(int ws, int sx, int sy);
hipError_t launch_search_i8(const SearchArgs &a, hipStream_t st);
bool search_mt_supported(int ws, int sx, int sy, int bits);
hipError_t launch_search_mt(const SearchArgs &a, hipStream_t st);
// q-factors (PCE, PPE) of every subset into a.qf (2S, normalised per column over the
// subsets), from the full NCC map; LDS need per workgroup (limit 150 KiB).
size_t qfactor_lds_bytes(int ws, int sx, int sy);
hipError_t launch_qfactors(const SearchArgs &a, hipStream_t st);

// Outlier removal and NaN in-painting on strided device views (outlier.hip): element
// (i, j) at p[i si + j sj].  Return 0, -1 (HIP error) or an ALDIC error code with msg.
OutlierWork *outlier_work_new();
void outlier_work_free(OutlierWork *w);

// Learned initial guess (cnn.hip): a siamese correlation network on the 64 x 64 patch
// pairs (plus a 4-pixel context ring) of the nodes.  All pointers are device pointers
// unless named _host; every call enqueues on st.
// Multigrid integer search (funIntegerSearchMg.m:3-285): final-level region centres and
// values, the fitted grid (xlist x ylist, x fastest) and the gridfit CG iterations.
struct MgResult {
    int levels = 0;
    std::vector<double> tempx, tempy, tempu, tempv, tempphi, xlist, ylist, u, v, cc;
    int fit_iters[2] = {0, 0};
};
MgWork *mg_work_new();
void mg_work_free(MgWork *w);
// 0 ok, 1 argument/data error (err says which), 2 HIP error
int integer_search_mg(MgWork &W, const float *f, const float *g, int nx, int ny, const int *grid, int ws,
                      int step, MgResult &res, std::string &err, hipStream_t st);
CnnWork *cnn_work_new();
void cnn_work_free(CnnWork *c);
long cnn_nparams();
hipError_t cnn_set_weights(CnnWork &c, const float *w_host, hipStream_t st);  // waits
hipError_t cnn_get_weights(CnnWork &c, float *w_host, hipStream_t st);        // waits
// How cnn_forward runs on a node set (host arithmetic on the nodes, 2 per node, 1-based):
// on one 4-pixel lattice of px0, py0 and nx1 x ny1 E1 cells, or per patch
CnnPlan cnn_plan(const double *xy_host, int S);
// P (6 per node: u_x, v_x, u_y, v_y, u, v) at nodes xy (2 per node, 1-based) from f, g
hipError_t cnn_forward(CnnWork &c, const float *f, const float *g, int nx, int ny,
                       const double *xy, const CnnPlan &plan, int S, double *P, hipStream_t st);
// appends the nodes' patch pairs and targets P (6 per node) to the training set
hipError_t cnn_add_samples(CnnWork &c, const float *f, const float *g, int nx, int ny,
                           const double *xy, const double *P, int S, hipStream_t st);
void cnn_clear_samples(CnnWork &c);
long cnn_n_samples(const CnnWork &c);
// forward + backward of training-set rows idx_host (B): loss to *loss_dev, gradient kept;
// one Adam step when lr > 0
hipError_t cnn_step(CnnWork &c, const int *idx_host, int B, float lr, double *loss_dev,
                    hipStream_t st);
const float *cnn_grad(const CnnWork &c);     // the last step's gradient (nparams)
bool cnn_have_weights(const CnnWork &c);
const float *cnn_outputs(const CnnWork &c);  // the last forward's scaled outputs (6 B)
// LocalICGN with the learned start (aldic_local_icgn): node centres of a search grid,
// the network's P split into U0 and F0, and the fallback flags
hipError_t launch_grid_coords(int x1, int y1, int Nx, int Ny, int step, int ws, double *xy, hipStream_t st);
hipError_t launch_split_start(const double *P, int S, double *U0, double *F0, hipStream_t st);
hipError_t launch_flag_fallback(const int *it, const double *z, int S, int max_it, double min_z, int *flag,
                                hipStream_t st);
int inpaint_view(OutlierWork &w, double *p, int rows, int cols, long si, long sj, int method,
                 hipStream_t st, const char *&msg);
int remove_outliers_views(OutlierWork &w, double *up, double *vp, int rows, int cols, long si,
                          long sj, float thr, bool keep_nan, int *flagged, long fsi, long fsj,
                          int *n_flagged, hipStream_t st, const char *&msg);
hipError_t launch_mask_nodes(double *U, const double *coords, int S, const uint8_t *mask, int nx,
                             int ny, hipStream_t st);

// Strain post-processing on a uniform M x N node grid (strain.hip): F from U by finite
// differences (method 1) or finite differences plus plane fits of half width R inside
// the R ring (method 2); then the strain type in place on all nodes and the trimmed,
// world-oriented copy W of the nodes at least r from the border.
hipError_t launch_strain_grad(int M, int N, double h, int method, int R, const double *U,
                              double *F, hipStream_t st);
hipError_t launch_strain_finish(int M, int N, int type, int r, double *F, double *W,
                                hipStream_t st);

hipError_t launch_bicubic(const float *g, int nx, int ny, const double *X, const double *Y,
                          long n, double *Z, hipStream_t st);

// FEM
hipError_t fem_assemble(MeshDev &m, double beta, double mu, double alpha, hipStream_t st);
hipError_t fem_rhs(const MeshDev &m, double beta, double mu, const double *U, const double *F,
                   const double *udual, const double *vdual, double *b, hipStream_t st);
// stats_dev (nullable): the device copy of {converged, iterations, residual} of the last
// iteration kernel (no host wait).  warm: x holds the initial guess on entry (else 0).
hipError_t fem_pcg(const MeshDev &m, const double *b, double *x, double rtol, int maxit,
                   DBuf *work, double *stats_host, hipStream_t st, int poll,
                   double *stats_dev = nullptr, bool warm = false);
// Solves the assembled system: fast diagonalisation (direct, FP64 GEMMs) on structured
// grids, Jacobi-PCG otherwise.  stats_host (nullable) = {iterations (0 = direct),
// relative residual estimate}.
hipError_t fem_solve(const MeshDev &m, const double *b, double *x, double rtol, int maxit,
                     DBuf *work, double *stats_host, hipStream_t st, int poll);
// Nodal strain; with udual set also udual = F - F1, vdual = U - U1 (ADMM dual update).
hipError_t fem_strain(const MeshDev &m, const double *U, int swap_fix, double *F, hipStream_t st,
                      const double *F1 = nullptr, const double *U1 = nullptr,
                      double *udual = nullptr, double *vdual = nullptr);

// Quadtree global step: assemble beta K + mu M with the fixed nodes' rows made identity
// rows; rhs b = beta G . F + mu M U at free nodes minus the fixed columns' share, 0 at
// fixed nodes; quad_finish writes the fixed nodes' values into the solution.
hipError_t quad_assemble(QuadDev &q, double beta, double mu, hipStream_t st);
hipError_t quad_rhs(QuadDev &q, double beta, double mu, const double *U, const double *F,
                    double *b, hipStream_t st);
hipError_t quad_finish(QuadDev &q, const double *U, double *x, hipStream_t st);
// x = U at the free nodes, 0 at the fixed ones (a warm start for the quadtree PCG)
hipError_t quad_start(const QuadDev &q, const double *U, double *x, hipStream_t st);
// Gauss-point strains of the quadtree mesh (midside nodes dropped): Fgp 4E x [u_x u_y v_x
// v_y], gpxy (nullable) 4E x 2; and node values by given linear-interpolation weights.
hipError_t quad_gp_strain(const QuadDev &q, const double *U, double *Fgp, double *gpxy,
                          hipStream_t st);
hipError_t quad_gp_gather(int nn, const int *tri, const double *w, const double *Fgp, double *F,
                          hipStream_t st);

// ADMM helpers
hipError_t launch_sub(const double *a, const double *b, double *out, long n, hipStream_t st);
// out[0] = sum (a - b)^2 (device scalar)
hipError_t launch_sqdiff(const double *a, const double *b, long n, double *out, hipStream_t st);
// nbad (nullable, device): the number of statuses < 0 or > 99 is written there.
hipError_t launch_badfill(const int *iters, int S, int Nx, int Ny, int use_mean_rule,
                          const double *Usrc, const double *Fsrc, double *U, double *F, int *bad,
                          double *part, hipStream_t st, int *nbad = nullptr);
// The bad-point rule alone (rule 0: status < 0 or > 99; 1: also > mean + 0.15 std of the
// good ones; 2: also > max(mean + 0.15 std, 20)); bad U (and F) become NaN; nbad as above.
hipError_t launch_badmark(const int *iters, int S, int rule, const double *Usrc,
                          const double *Fsrc, double *U, double *F, int *bad, double *part,
                          hipStream_t st, int *nbad = nullptr);
// p[0 .. n) = v (device ints).
hipError_t launch_iset(int *p, int n, int v, hipStream_t st);
// Refill of the bad nodes of an unstructured mesh from the node adjacency (rowptr/col):
// sweeps over the nodes still unknown, each taking the inverse-distance average of its
// known neighbours; nodes left unknown after `sweeps` get 0.  known: int[nn] scratch.
hipError_t launch_meshfill(const MeshDev &m, const int *bad, double *U, double *F, int *known,
                           int sweeps, hipStream_t st);
// x[i] = src[i] at the Dirichlet-coded nodes of the quadtree (4 values per node)
hipError_t quad_copy_dirichlet(const QuadDev &q, const double *src, double *x, hipStream_t st);

}  // namespace aldic
