// aldic_mex.cpp -- MATLAB MEX gateway of the AL-DIC library (include/aldic.h).
//
// Build (MATLAB R2018a or later, interleaved complex API), from this directory:
//   mex -R2018a -I../../include aldic_mex.cpp -L../lib -laldic
// and put ../lib on LD_LIBRARY_PATH.  One gateway, one library context per MATLAB session
// (created on the first call, destroyed at exit), commands selected by the first argument:
//   aldic_mex('command', ...).  The .m wrappers in this directory give the reference's
// functions (IntegerSearch, LocalICGN, Subpb1, Subpb2, funGlobal_NodalStrainAvg and the
// quadtree variants) their exact signatures on top of these commands.
//
// Conventions (as the reference's arrays arrive from MATLAB):
//   * images are MATLAB's transposed images (ReadImage.m:79): size nx x ny column-major,
//     which is the library's row-major [y][x] memory, passed through without a copy;
//   * coordinate lists are S x 2 column-major and are interleaved here;
//   * element lists are 1-based (0 = absent midside node) and are made 0-based here;
//   * U, F, HPar vectors keep the reference's interleaving.
// Errors raise a MATLAB error carrying aldic_last_error().
//
// Exports of aldic.h not reachable from MATLAB arrays (no command): aldic_create /
// aldic_destroy / aldic_last_error (managed by this gateway) and aldic_prep_noalloc (takes
// device pointers).
#include <cstring>
#include <string>
#include <vector>

#include "aldic.h"
#include "mex.h"

namespace {

aldic_ctx *g_ctx = nullptr;

void cleanup()
{
    aldic_destroy(g_ctx);
    g_ctx = nullptr;
}

void check(int rc, const char *what)
{
    if (rc) mexErrMsgIdAndTxt("aldic:call", "%s: %s", what, aldic_last_error(g_ctx));
}

void need(int nrhs, int n, const char *cmd)
{
    if (nrhs < n) mexErrMsgIdAndTxt("aldic:args", "%s: %d arguments expected", cmd, n - 1);
}

const double *dbl(const mxArray *a, const char *what)
{
    if (!mxIsDouble(a) || mxIsComplex(a)) mexErrMsgIdAndTxt("aldic:args", "%s must be real double", what);
    return mxGetDoubles(a);
}

size_t numel(const mxArray *a) { return mxGetNumberOfElements(a); }

// S x 2 column-major -> interleaved
std::vector<double> coords2(const mxArray *a)
{
    const size_t S = mxGetM(a);
    const double *c = dbl(a, "coordinates");
    if (mxGetN(a) != 2) mexErrMsgIdAndTxt("aldic:args", "coordinates must be S x 2");
    std::vector<double> xy(2 * S);
    for (size_t j = 0; j < S; ++j) { xy[2 * j] = c[j]; xy[2 * j + 1] = c[S + j]; }
    return xy;
}

std::vector<int> ints(const mxArray *a, int offset)
{
    const size_t n = numel(a);
    std::vector<int> v(n);
    const double *p = dbl(a, "index list");
    for (size_t k = 0; k < n; ++k) v[k] = (int)p[k] + offset;
    return v;
}

// E x ncol 1-based element list (0 = absent) -> 0-based, ncol_out columns, -1 = absent
std::vector<int> elems(const mxArray *a, int ncol_out)
{
    const size_t E = mxGetM(a), ncol = mxGetN(a);
    const double *p = dbl(a, "elements");
    std::vector<int> e((size_t)ncol_out * E, -1);
    for (size_t k = 0; k < E; ++k)
        for (size_t c = 0; c < ncol && (int)c < ncol_out; ++c)
            e[(size_t)ncol_out * k + c] = (int)p[c * E + k] - 1;
    return e;
}

int dtype_of(const mxArray *a)
{
    switch (mxGetClassID(a)) {
    case mxUINT8_CLASS: return ALDIC_U8;
    case mxUINT16_CLASS: return ALDIC_U16;
    case mxSINGLE_CLASS: return ALDIC_F32;
    case mxDOUBLE_CLASS: return ALDIC_F64;
    default: mexErrMsgIdAndTxt("aldic:args", "images must be uint8, uint16, single or double");
    }
    return -1;
}

mxArray *dmat(size_t m, size_t n) { return mxCreateDoubleMatrix(m, n, mxREAL); }

mxArray *imat(size_t m, size_t n) { return mxCreateNumericMatrix(m, n, mxINT32_CLASS, mxREAL); }

int *iptr(mxArray *a) { return static_cast<int *>(mxGetData(a)); }

// Nx*Ny row-major (x fastest) -> Ny x Nx MATLAB array (the reference's x, y, u, v, cc.max)
mxArray *grid_out(const std::vector<double> &v, int Nx, int Ny)
{
    mxArray *m = dmat(Ny, Nx);
    double *d = mxGetDoubles(m);
    for (int l = 0; l < Ny; ++l)
        for (int k = 0; k < Nx; ++k) d[(size_t)k * Ny + l] = v[(size_t)l * Nx + k];
    return m;
}

}  // namespace

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgIdAndTxt("aldic:args", "aldic_mex('command', ...)");
    if (!g_ctx) {
        if (aldic_create(&g_ctx, 0)) mexErrMsgIdAndTxt("aldic:create", "aldic_create failed");
        mexAtExit(cleanup);
    }
    char cbuf[64];
    mxGetString(prhs[0], cbuf, sizeof cbuf);
    const std::string cmd(cbuf);
    (void)nlhs;

    if (cmd == "set_images") {  // aldic_mex('set_images', f, g [, mask, roi])
        need(nrhs, 3, "set_images");
        const int nx = (int)mxGetM(prhs[1]), ny = (int)mxGetN(prhs[1]);
        std::vector<uint8_t> mask;
        if (nrhs > 3 && numel(prhs[3]) == (size_t)nx * ny) {
            const mxArray *m = prhs[3];
            mask.resize(numel(m));
            if (mxIsLogical(m)) {
                const mxLogical *lp = mxGetLogicals(m);
                for (size_t k = 0; k < mask.size(); ++k) mask[k] = lp[k] ? 1 : 0;
            } else {
                const double *dp = dbl(m, "mask");
                for (size_t k = 0; k < mask.size(); ++k) mask[k] = dp[k] != 0 ? 1 : 0;
            }
        }
        std::vector<int> roi;
        if (nrhs > 4 && numel(prhs[4]) == 4) roi = ints(prhs[4], 0);
        check(aldic_set_images(g_ctx, mxGetData(prhs[1]), mxGetData(prhs[2]), nx, ny,
                               dtype_of(prhs[1]), mask.empty() ? nullptr : mask.data(),
                               roi.empty() ? nullptr : roi.data()),
              "set_images");
    } else if (cmd == "set_normalized") {  // aldic_mex('set_normalized', fNormalized, gNormalized)
        need(nrhs, 3, cmd.c_str());
        check(aldic_set_normalized(g_ctx, dbl(prhs[1], "f"), dbl(prhs[2], "g"), (int)mxGetM(prhs[1]),
                                   (int)mxGetN(prhs[1])),
              "set_normalized");
    } else if (cmd == "set_gradients") {  // aldic_mex('set_gradients', DfDx, DfDy): (nx-6) x (ny-6)
        need(nrhs, 3, cmd.c_str());
        const int mx = (int)mxGetM(prhs[1]), my = (int)mxGetN(prhs[1]);
        const int nx = mx + 6, ny = my + 6;
        std::vector<double> gx((size_t)nx * ny, 0.0), gy((size_t)nx * ny, 0.0);
        const double *a = dbl(prhs[1], "DfDx"), *b = dbl(prhs[2], "DfDy");
        for (int j = 0; j < my; ++j)
            for (int i = 0; i < mx; ++i) {  // DfDx(i, j) is the gradient at pixel (i+3, j+3)
                gx[(size_t)(j + 3) * nx + i + 3] = a[(size_t)j * mx + i];
                gy[(size_t)(j + 3) * nx + i + 3] = b[(size_t)j * mx + i];
            }
        check(aldic_set_gradients(g_ctx, gx.data(), gy.data()), "set_gradients");
    } else if (cmd == "get_images") {  // [f, g, gx, gy] = aldic_mex('get_images', nx, ny)
        need(nrhs, 3, cmd.c_str());
        const size_t nx = (size_t)mxGetScalar(prhs[1]), ny = (size_t)mxGetScalar(prhs[2]);
        for (int k = 0; k < 4; ++k) plhs[k] = dmat(nx, ny);
        check(aldic_get_images(g_ctx, mxGetDoubles(plhs[0]), mxGetDoubles(plhs[1]),
                               mxGetDoubles(plhs[2]), mxGetDoubles(plhs[3])),
              "get_images");
    } else if (cmd == "search_grid") {  // out = aldic_mex('search_grid', nx, ny, grid, ws, step, s)
        need(nrhs, 7, cmd.c_str());
        const std::vector<int> grid = ints(prhs[3], 0), s = ints(prhs[6], 0);
        int out[4];
        if (aldic_search_grid((int)mxGetScalar(prhs[1]), (int)mxGetScalar(prhs[2]), grid.data(),
                              (int)mxGetScalar(prhs[4]), (int)mxGetScalar(prhs[5]), s.data(), out))
            mexErrMsgIdAndTxt("aldic:args", "search_grid: bad arguments");
        plhs[0] = dmat(1, 4);
        for (int k = 0; k < 4; ++k) mxGetDoubles(plhs[0])[k] = out[k];
    } else if (cmd == "integer_search") {
        // [x, y, u, v, ccmax, qf, status] = aldic_mex('integer_search', nx, ny, grid, ws, step, s, want_qf)
        need(nrhs, 8, cmd.c_str());
        const std::vector<int> grid = ints(prhs[3], 0), s = ints(prhs[6], 0);
        const int ws = (int)mxGetScalar(prhs[4]), step = (int)mxGetScalar(prhs[5]);
        int g4[4];
        if (aldic_search_grid((int)mxGetScalar(prhs[1]), (int)mxGetScalar(prhs[2]), grid.data(), ws,
                              step, s.data(), g4))
            mexErrMsgIdAndTxt("aldic:args", "integer_search: bad grid");
        const int Nx = g4[2], Ny = g4[3], S = Nx * Ny;
        const bool want_qf = nrhs > 7 && mxGetScalar(prhs[7]) != 0;
        std::vector<double> x(S), y(S), u(S), v(S), cm(S), qf(want_qf ? 2 * (size_t)S : 0);
        std::vector<int> st(S);
        check(aldic_integer_search(g_ctx, grid.data(), ws, step, s.data(), x.data(), y.data(),
                                   u.data(), v.data(), cm.data(), want_qf ? qf.data() : nullptr,
                                   st.data()),
              "integer_search");
        plhs[0] = grid_out(x, Nx, Ny);
        plhs[1] = grid_out(y, Nx, Ny);
        plhs[2] = grid_out(u, Nx, Ny);
        plhs[3] = grid_out(v, Nx, Ny);
        plhs[4] = grid_out(cm, Nx, Ny);
        plhs[5] = dmat(want_qf ? S : 0, 2);  // cc.qfactors, S x 2
        for (int q = 0; want_qf && q < S; ++q) {
            mxGetDoubles(plhs[5])[q] = qf[2 * (size_t)q];
            mxGetDoubles(plhs[5])[S + q] = qf[2 * (size_t)q + 1];
        }
        plhs[6] = imat(S, 1);
        std::memcpy(iptr(plhs[6]), st.data(), sizeof(int) * S);
    } else if (cmd == "integer_search_mg") {
        // [x, y, u, v, ccmax, dims] = aldic_mex('integer_search_mg', grid, ws, step)
        need(nrhs, 4, cmd.c_str());
        const std::vector<int> grid = ints(prhs[1], 0);
        if (grid.size() != 4) mexErrMsgIdAndTxt("aldic:args", "integer_search_mg: grid = [x1 x2 y1 y2]");
        const int ws = (int)mxGetScalar(prhs[2]), step = (int)mxGetScalar(prhs[3]);
        if (ws < 1 || step < 1) mexErrMsgIdAndTxt("aldic:args", "integer_search_mg: bad ws / step");
        const long cap = std::max(4L, (long)((grid[1] - grid[0]) / step + 1) * ((grid[3] - grid[2]) / step + 1));
        std::vector<double> x(cap), y(cap), u(cap), v(cap), cm(cap);
        int dims[6];
        check(aldic_integer_search_mg(g_ctx, grid.data(), ws, step, cap, dims, x.data(), y.data(), u.data(),
                                      v.data(), cm.data(), 0, nullptr),
              "integer_search_mg");
        const int Nx = dims[0], Ny = dims[1];
        plhs[0] = grid_out(x, Nx, Ny);
        plhs[1] = grid_out(y, Nx, Ny);
        plhs[2] = grid_out(u, Nx, Ny);
        plhs[3] = grid_out(v, Nx, Ny);
        plhs[4] = grid_out(cm, Nx, Ny);
        plhs[5] = imat(6, 1);
        std::memcpy(iptr(plhs[5]), dims, sizeof(dims));
    } else if (cmd == "icgn6") {
        // [U, F, H21, iters, zncc] = aldic_mex('icgn6', U0, coords, ws, tol): ws scalar or S-vector
        need(nrhs, 5, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[2]);
        const int S = (int)(xy.size() / 2);
        const bool per = numel(prhs[3]) == (size_t)S && S > 1;
        const std::vector<int> wsv = per ? ints(prhs[3], 0) : std::vector<int>();
        plhs[0] = dmat(2 * (size_t)S, 1);
        plhs[1] = dmat(4 * (size_t)S, 1);
        plhs[2] = dmat(21, S);  // the .m wrapper transposes to the reference's S x 21
        plhs[3] = imat(S, 1);
        plhs[4] = dmat(S, 1);
        check(aldic_icgn6(g_ctx, S, xy.data(), dbl(prhs[1], "U0"), per ? wsv.data() : nullptr,
                          per ? 0 : (int)mxGetScalar(prhs[3]), mxGetScalar(prhs[4]),
                          mxGetDoubles(plhs[0]), mxGetDoubles(plhs[1]), mxGetDoubles(plhs[2]),
                          iptr(plhs[3]), mxGetDoubles(plhs[4])),
              "icgn6");
    } else if (cmd == "icgn6_start") {
        // [U, F, H21, iters, zncc] = aldic_mex('icgn6_start', P0 (6 x S), coords, ws, tol)
        need(nrhs, 5, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[2]);
        const int S = (int)(xy.size() / 2);
        if (numel(prhs[1]) != 6 * (size_t)S) mexErrMsgIdAndTxt("aldic:args", "P0 must be 6 x S");
        const bool per = numel(prhs[3]) == (size_t)S && S > 1;
        const std::vector<int> wsv = per ? ints(prhs[3], 0) : std::vector<int>();
        plhs[0] = dmat(2 * (size_t)S, 1);
        plhs[1] = dmat(4 * (size_t)S, 1);
        plhs[2] = dmat(21, S);
        plhs[3] = imat(S, 1);
        plhs[4] = dmat(S, 1);
        check(aldic_icgn6_start(g_ctx, S, xy.data(), dbl(prhs[1], "P0"), per ? wsv.data() : nullptr,
                                per ? 0 : (int)mxGetScalar(prhs[3]), mxGetScalar(prhs[4]),
                                mxGetDoubles(plhs[0]), mxGetDoubles(plhs[1]), mxGetDoubles(plhs[2]),
                                iptr(plhs[3]), mxGetDoubles(plhs[4])),
              "icgn6_start");
    } else if (cmd == "cnn_set_weights") {  // aldic_mex('cnn_set_weights', w) (single, nparams)
        need(nrhs, 2, cmd.c_str());
        if (mxGetClassID(prhs[1]) != mxSINGLE_CLASS || numel(prhs[1]) != (size_t)aldic_cnn_nparams())
            mexErrMsgIdAndTxt("aldic:args", "weights must be single with aldic_cnn_nparams() elements");
        check(aldic_cnn_set_weights(g_ctx, static_cast<const float *>(mxGetData(prhs[1])),
                                    aldic_cnn_nparams()),
              "cnn_set_weights");
    } else if (cmd == "cnn_get_weights") {  // w = aldic_mex('cnn_get_weights')
        const long n = aldic_cnn_nparams();
        plhs[0] = mxCreateNumericMatrix((size_t)n, 1, mxSINGLE_CLASS, mxREAL);
        check(aldic_cnn_get_weights(g_ctx, static_cast<float *>(mxGetData(plhs[0])), n), "cnn_get_weights");
    } else if (cmd == "cnn_predict") {  // P = aldic_mex('cnn_predict', coords): 6 x S
        need(nrhs, 2, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[1]);
        const int S = (int)(xy.size() / 2);
        plhs[0] = dmat(6, S);
        check(aldic_cnn_predict(g_ctx, S, xy.data(), mxGetDoubles(plhs[0])), "cnn_predict");
    } else if (cmd == "cnn_add_samples") {  // aldic_mex('cnn_add_samples', coords, P (6 x S))
        need(nrhs, 3, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[1]);
        const int S = (int)(xy.size() / 2);
        if (numel(prhs[2]) != 6 * (size_t)S) mexErrMsgIdAndTxt("aldic:args", "P must be 6 x S");
        check(aldic_cnn_add_samples(g_ctx, S, xy.data(), dbl(prhs[2], "P")), "cnn_add_samples");
    } else if (cmd == "cnn_clear_samples") {  // aldic_mex('cnn_clear_samples')
        check(aldic_cnn_clear_samples(g_ctx), "cnn_clear_samples");
    } else if (cmd == "cnn_train") {  // loss = aldic_mex('cnn_train', steps, batch, lr, seed)
        need(nrhs, 5, cmd.c_str());
        const int steps = (int)mxGetScalar(prhs[1]);
        plhs[0] = dmat(steps > 0 ? (size_t)steps : 0, 1);
        check(aldic_cnn_train(g_ctx, steps, (int)mxGetScalar(prhs[2]), (float)mxGetScalar(prhs[3]),
                              (unsigned long long)mxGetScalar(prhs[4]), mxGetDoubles(plhs[0])),
              "cnn_train");
    } else if (cmd == "cnn_grad") {  // [g, loss, out] = aldic_mex('cnn_grad', rows (0-based))
        need(nrhs, 2, cmd.c_str());
        const std::vector<int> rows = ints(prhs[1], 0);
        const int B = (int)rows.size();
        plhs[0] = mxCreateNumericMatrix((size_t)aldic_cnn_nparams(), 1, mxSINGLE_CLASS, mxREAL);
        plhs[1] = dmat(1, 1);
        plhs[2] = dmat(6, (size_t)B);
        check(aldic_cnn_grad(g_ctx, B, rows.data(), static_cast<float *>(mxGetData(plhs[0])),
                             mxGetDoubles(plhs[1]), mxGetDoubles(plhs[2])),
              "cnn_grad");
    } else if (cmd == "local_icgn") {
        // [x, y, U, F, iters, zncc, fallback, ms] = aldic_mex('local_icgn', nx, ny, start, grid, ws,
        //     step, s, tol, max_it, min_zncc)
        need(nrhs, 11, cmd.c_str());
        const std::vector<int> grid = ints(prhs[4], 0), s = ints(prhs[7], 0);
        const int ws = (int)mxGetScalar(prhs[5]), step = (int)mxGetScalar(prhs[6]);
        int g4[4];
        if (aldic_search_grid((int)mxGetScalar(prhs[1]), (int)mxGetScalar(prhs[2]), grid.data(), ws,
                              step, s.data(), g4))
            mexErrMsgIdAndTxt("aldic:args", "local_icgn: bad grid");
        const int S = g4[2] * g4[3];
        plhs[0] = dmat(S, 1);
        plhs[1] = dmat(S, 1);
        plhs[2] = dmat(2 * (size_t)S, 1);
        plhs[3] = dmat(4 * (size_t)S, 1);
        plhs[4] = imat(S, 1);
        plhs[5] = dmat(S, 1);
        plhs[6] = imat(S, 1);
        std::vector<float> ms(5);
        int nfb = 0;
        check(aldic_local_icgn(g_ctx, (int)mxGetScalar(prhs[3]), grid.data(), ws, step, s.data(),
                               mxGetScalar(prhs[8]), (int)mxGetScalar(prhs[9]), mxGetScalar(prhs[10]),
                               mxGetDoubles(plhs[0]), mxGetDoubles(plhs[1]), mxGetDoubles(plhs[2]),
                               mxGetDoubles(plhs[3]), iptr(plhs[4]), mxGetDoubles(plhs[5]), iptr(plhs[6]),
                               &nfb, ms.data()),
              "local_icgn");
        plhs[7] = dmat(1, 5);
        for (int k = 0; k < 5; ++k) mxGetDoubles(plhs[7])[k] = ms[k];
    } else if (cmd == "icgn2") {
        // [U, iters] = aldic_mex('icgn2', coords, ws, H21 (21 x S), mu, vdual, Uold, Fold, tol)
        need(nrhs, 9, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[1]);
        const int S = (int)(xy.size() / 2);
        const bool per = numel(prhs[2]) == (size_t)S && S > 1;
        const std::vector<int> wsv = per ? ints(prhs[2], 0) : std::vector<int>();
        plhs[0] = dmat(2 * (size_t)S, 1);
        plhs[1] = imat(S, 1);
        check(aldic_icgn2(g_ctx, S, xy.data(), per ? wsv.data() : nullptr,
                          per ? 0 : (int)mxGetScalar(prhs[2]), dbl(prhs[3], "H21"),
                          mxGetScalar(prhs[4]), dbl(prhs[5], "vdual"), dbl(prhs[6], "Uold"),
                          dbl(prhs[7], "Fold"), mxGetScalar(prhs[8]), mxGetDoubles(plhs[0]),
                          iptr(plhs[1])),
              "icgn2");
    } else if (cmd == "set_mesh") {  // aldic_mex('set_mesh', coordinatesFEM, elementsFEM)
        need(nrhs, 3, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[1]);
        const std::vector<int> e = elems(prhs[2], 4);
        check(aldic_set_mesh(g_ctx, xy.data(), (int)(xy.size() / 2), e.data(), (int)(e.size() / 4)),
              "set_mesh");
    } else if (cmd == "subpb2") {
        // [Uhat, stats] = aldic_mex('subpb2', beta, mu, alpha, U, F, udual, vdual, rtol, maxit)
        need(nrhs, 10, cmd.c_str());
        plhs[0] = dmat(numel(prhs[4]), 1);
        plhs[1] = dmat(1, 2);
        check(aldic_subpb2(g_ctx, mxGetScalar(prhs[1]), mxGetScalar(prhs[2]), mxGetScalar(prhs[3]),
                           dbl(prhs[4], "U"), dbl(prhs[5], "F"), dbl(prhs[6], "udual"),
                           dbl(prhs[7], "vdual"), mxGetScalar(prhs[8]), (int)mxGetScalar(prhs[9]),
                           mxGetDoubles(plhs[0]), mxGetDoubles(plhs[1])),
              "subpb2");
    } else if (cmd == "nodal_strain") {  // F = aldic_mex('nodal_strain', U, swap_fix)
        need(nrhs, 3, cmd.c_str());
        plhs[0] = dmat(2 * numel(prhs[1]), 1);
        check(aldic_nodal_strain(g_ctx, dbl(prhs[1], "U"), (int)mxGetScalar(prhs[2]),
                                 mxGetDoubles(plhs[0])),
              "nodal_strain");
    } else if (cmd == "inpaint_nans") {  // B = aldic_mex('inpaint_nans', A, method)
        need(nrhs, 3, cmd.c_str());
        plhs[0] = mxDuplicateArray(prhs[1]);
        check(aldic_inpaint_nans(g_ctx, mxGetDoubles(plhs[0]), (int)mxGetM(prhs[1]),
                                 (int)mxGetN(prhs[1]), (int)mxGetScalar(prhs[2])),
              "inpaint_nans");
    } else if (cmd == "remove_outliers") {
        // [u, v, flagged] = aldic_mex('remove_outliers', u, v, Thr0)
        need(nrhs, 4, cmd.c_str());
        const int rows = (int)mxGetM(prhs[1]), cols = (int)mxGetN(prhs[1]);
        plhs[0] = mxDuplicateArray(prhs[1]);
        plhs[1] = mxDuplicateArray(prhs[2]);
        std::vector<int> fl((size_t)rows * cols);
        int nf = 0;
        check(aldic_remove_outliers(g_ctx, mxGetDoubles(plhs[0]), mxGetDoubles(plhs[1]), rows, cols,
                                    mxGetScalar(prhs[3]), fl.data(), &nf),
              "remove_outliers");
        plhs[2] = dmat(rows, cols);
        for (size_t k = 0; k < fl.size(); ++k) mxGetDoubles(plhs[2])[k] = fl[k];
    } else if (cmd == "compute_strain") {
        // [Ffinite, Fworld] = aldic_mex('compute_strain', M, N, winstepsize, ULocal, FLocal,
        //                               MethodToComputeStrain, Rad, StrainType)
        need(nrhs, 9, cmd.c_str());
        const int M = (int)mxGetScalar(prhs[1]), N = (int)mxGetScalar(prhs[2]);
        const int method = (int)mxGetScalar(prhs[6]), rad = (int)mxGetScalar(prhs[7]);
        const int r = method == 0 ? 0 : (method == 2 ? rad : 1);
        const int Mi = M - 2 * r > 0 ? M - 2 * r : 0, Ni = N - 2 * r > 0 ? N - 2 * r : 0;
        plhs[0] = dmat(4 * (size_t)M * N, 1);
        plhs[1] = dmat(4 * (size_t)Mi * Ni, 1);
        const double *U = numel(prhs[4]) ? dbl(prhs[4], "ULocal") : nullptr;
        const double *F = numel(prhs[5]) ? dbl(prhs[5], "FLocal") : nullptr;
        check(aldic_compute_strain(g_ctx, M, N, mxGetScalar(prhs[3]), U, F, method, rad,
                                   (int)mxGetScalar(prhs[8]), mxGetDoubles(plhs[0]),
                                   (Mi > 0 && Ni > 0) ? mxGetDoubles(plhs[1]) : nullptr),
              "compute_strain");
    } else if (cmd == "set_quadtree") {  // aldic_mex('set_quadtree', coords, elementsFEM (E x 8), dirichlet)
        need(nrhs, 4, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[1]);
        const std::vector<int> e = elems(prhs[2], 8), d = ints(prhs[3], -1);
        check(aldic_set_quadtree(g_ctx, xy.data(), (int)(xy.size() / 2), e.data(),
                                 (int)(e.size() / 8), d.empty() ? nullptr : d.data(), (int)d.size()),
              "set_quadtree");
    } else if (cmd == "subpb2_quadtree") {
        // [Uhat, stats] = aldic_mex('subpb2_quadtree', beta, mu, U, F, rtol, maxit)
        need(nrhs, 7, cmd.c_str());
        plhs[0] = dmat(numel(prhs[3]), 1);
        plhs[1] = dmat(1, 2);
        check(aldic_subpb2_quadtree(g_ctx, mxGetScalar(prhs[1]), mxGetScalar(prhs[2]),
                                    dbl(prhs[3], "U"), dbl(prhs[4], "F"), mxGetScalar(prhs[5]),
                                    (int)mxGetScalar(prhs[6]), mxGetDoubles(plhs[0]),
                                    mxGetDoubles(plhs[1])),
              "subpb2_quadtree");
    } else if (cmd == "nodal_strain_quadtree") {
        // [F, Fgp, gpxy] = aldic_mex('nodal_strain_quadtree', U, E, tri (1-based, nn x 3), w)
        need(nrhs, 3, cmd.c_str());
        const size_t nn = numel(prhs[1]) / 2, E = (size_t)mxGetScalar(prhs[2]);
        const bool weights = nrhs > 4 && numel(prhs[3]) == 3 * nn;
        std::vector<int> tri;
        std::vector<double> w;
        if (weights) {  // nn x 3 column-major -> row-major, 0-based
            const double *t = dbl(prhs[3], "tri"), *ww = dbl(prhs[4], "w");
            tri.resize(3 * nn);
            w.resize(3 * nn);
            for (size_t i = 0; i < nn; ++i)
                for (int k = 0; k < 3; ++k) {
                    tri[3 * i + k] = (int)t[k * nn + i] - 1;
                    w[3 * i + k] = ww[k * nn + i];
                }
        }
        plhs[0] = dmat(weights ? 4 * nn : 0, 1);
        std::vector<double> Fgp(16 * E), gxy(8 * E);
        check(aldic_nodal_strain_quadtree(g_ctx, dbl(prhs[1], "U"), weights ? tri.data() : nullptr,
                                          weights ? w.data() : nullptr,
                                          weights ? mxGetDoubles(plhs[0]) : nullptr, Fgp.data(),
                                          gxy.data()),
              "nodal_strain_quadtree");
        plhs[1] = dmat(4 * E, 4);
        plhs[2] = dmat(4 * E, 2);
        for (size_t r = 0; r < 4 * E; ++r) {
            for (int c = 0; c < 4; ++c) mxGetDoubles(plhs[1])[c * 4 * E + r] = Fgp[4 * r + c];
            for (int c = 0; c < 2; ++c) mxGetDoubles(plhs[2])[c * 4 * E + r] = gxy[2 * r + c];
        }
    } else if (cmd == "natural_fill") {
        // [U, F] = aldic_mex('natural_fill', coords, bad (logical / 0-1), U, F)
        need(nrhs, 5, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[1]);
        const size_t nn = xy.size() / 2;
        std::vector<int> bad(nn);
        for (size_t i = 0; i < nn; ++i)
            bad[i] = mxIsLogical(prhs[2]) ? (mxGetLogicals(prhs[2])[i] ? 1 : 0)
                                           : (dbl(prhs[2], "bad")[i] != 0 ? 1 : 0);
        plhs[0] = mxDuplicateArray(prhs[3]);
        const bool withF = numel(prhs[4]) == 4 * nn;
        plhs[1] = withF ? mxDuplicateArray(prhs[4]) : dmat(0, 1);
        check(aldic_natural_fill(g_ctx, (int)nn, xy.data(), bad.data(), mxGetDoubles(plhs[0]),
                                 withF ? mxGetDoubles(plhs[1]) : nullptr),
              "natural_fill");
    } else if (cmd == "smooth_strain_quadtree") {
        // F = aldic_mex('smooth_strain_quadtree', F, coords, h, smoothness, filter_std, filter_size)
        need(nrhs, 7, cmd.c_str());
        const std::vector<double> xy = coords2(prhs[2]);
        const size_t nn = xy.size() / 2;
        if (numel(prhs[1]) != 4 * nn) mexErrMsgIdAndTxt("aldic:args", "F must hold 4 values per node");
        plhs[0] = mxDuplicateArray(prhs[1]);
        check(aldic_smooth_strain_quadtree(g_ctx, (int)nn, xy.data(), mxGetScalar(prhs[3]), mxGetScalar(prhs[4]),
                                           mxGetScalar(prhs[5]), (int)mxGetScalar(prhs[6]),
                                           mxGetDoubles(plhs[0]), nullptr),
              "smooth_strain_quadtree");
    } else if (cmd == "bicubic") {  // Z = aldic_mex('bicubic', X, Y): ba_interp2 of the deformed image
        need(nrhs, 3, cmd.c_str());
        const size_t n = numel(prhs[1]);
        plhs[0] = mxCreateNumericArray(mxGetNumberOfDimensions(prhs[1]), mxGetDimensions(prhs[1]),
                                       mxDOUBLE_CLASS, mxREAL);
        check(aldic_bicubic(g_ctx, dbl(prhs[1], "X"), dbl(prhs[2], "Y"), (long)n,
                            mxGetDoubles(plhs[0])),
              "bicubic");
    } else if (cmd == "gridfit") {
        // zgrid = aldic_mex('gridfit', xnodes, ynodes, z (gy x gx, NaN = no datum), rtol, maxit)
        need(nrhs, 6, cmd.c_str());
        const int gx = (int)numel(prhs[1]), gy = (int)numel(prhs[2]);
        // MATLAB's zgrid(iy, ix) is column-major; the library's is [iy][ix] row-major
        std::vector<double> z((size_t)gx * gy), out((size_t)gx * gy);
        const double *zm = dbl(prhs[3], "z");
        for (int ix = 0; ix < gx; ++ix)
            for (int iy = 0; iy < gy; ++iy) z[(size_t)iy * gx + ix] = zm[(size_t)ix * gy + iy];
        double st[2];
        check(aldic_gridfit(g_ctx, gx, gy, dbl(prhs[1], "xnodes"), dbl(prhs[2], "ynodes"), z.data(),
                            mxGetScalar(prhs[4]), (int)mxGetScalar(prhs[5]), out.data(), st),
              "gridfit");
        plhs[0] = dmat(gy, gx);
        for (int ix = 0; ix < gx; ++ix)
            for (int iy = 0; iy < gy; ++iy)
                mxGetDoubles(plhs[0])[(size_t)ix * gy + iy] = out[(size_t)iy * gx + ix];
    } else if (cmd == "frame_stage") {  // aldic_mex('frame_stage', f, g [, mask, roi])
        need(nrhs, 3, cmd.c_str());
        const int nx = (int)mxGetM(prhs[1]), ny = (int)mxGetN(prhs[1]);
        std::vector<int> roi;
        if (nrhs > 4 && numel(prhs[4]) == 4) roi = ints(prhs[4], 0);
        check(aldic_frame_stage(g_ctx, mxGetData(prhs[1]), mxGetData(prhs[2]), nx, ny,
                                dtype_of(prhs[1]), nullptr, roi.empty() ? nullptr : roi.data()),
              "frame_stage");
    } else if (cmd == "frame_setup") {
        // S = aldic_mex('frame_setup', ws, step, s, grid, tol, mu, beta, n_admm, cg_maxit,
        //               cg_rtol, smooth, beta_sweep, admm_tol [, outlier_thr (100; 0: off)])
        need(nrhs, 14, cmd.c_str());
        aldic_frame_opts o{};
        o.ws = (int)mxGetScalar(prhs[1]);
        o.step = (int)mxGetScalar(prhs[2]);
        const std::vector<int> s = ints(prhs[3], 0), grid = ints(prhs[4], 0);
        o.s[0] = s[0];
        o.s[1] = s.size() > 1 ? s[1] : s[0];
        for (int k = 0; k < 4; ++k) o.grid[k] = grid[k];
        o.tol = mxGetScalar(prhs[5]);
        o.mu = mxGetScalar(prhs[6]);
        o.beta = mxGetScalar(prhs[7]);
        o.n_admm = (int)mxGetScalar(prhs[8]);
        o.cg_maxit = (int)mxGetScalar(prhs[9]);
        o.cg_rtol = mxGetScalar(prhs[10]);
        o.smooth = (int)mxGetScalar(prhs[11]);
        o.beta_sweep = (int)mxGetScalar(prhs[12]);
        o.admm_tol = mxGetScalar(prhs[13]);
        o.outlier_thr = nrhs > 14 ? mxGetScalar(prhs[14]) : 100.0;
        o.learned_start = nrhs > 15 ? (int)mxGetScalar(prhs[15]) : 0;
        int S = 0;
        check(aldic_frame_setup(g_ctx, &o, &S), "frame_setup");
        plhs[0] = mxCreateDoubleScalar(S);
    } else if (cmd == "frame_run") {  // aldic_mex('frame_run')
        check(aldic_frame_run(g_ctx), "frame_run");
    } else if (cmd == "frame_run_seeded") {  // aldic_mex('frame_run_seeded', U0): previous frame's U
        need(nrhs, 2, cmd.c_str());
        check(aldic_frame_run_seeded(g_ctx, dbl(prhs[1], "U0")), "frame_run_seeded");
    } else if (cmd == "frame_get") {  // [U, F, U6, F6, it6, it2, U0] = aldic_mex('frame_get', S)
        need(nrhs, 2, cmd.c_str());
        const size_t S = (size_t)mxGetScalar(prhs[1]);
        plhs[0] = dmat(2 * S, 1); plhs[1] = dmat(4 * S, 1); plhs[2] = dmat(2 * S, 1);
        plhs[3] = dmat(4 * S, 1); plhs[4] = imat(S, 1); plhs[5] = imat(S, 1); plhs[6] = dmat(2 * S, 1);
        check(aldic_frame_get(g_ctx, mxGetDoubles(plhs[0]), mxGetDoubles(plhs[1]),
                              mxGetDoubles(plhs[2]), mxGetDoubles(plhs[3]), iptr(plhs[4]),
                              iptr(plhs[5]), mxGetDoubles(plhs[6])),
              "frame_get");
    } else if (cmd == "frame_times") {  // ms = aldic_mex('frame_times')
        plhs[0] = dmat(1, 10);
        check(aldic_frame_times(g_ctx, mxGetDoubles(plhs[0]), 10), "frame_times");
    } else if (cmd == "frame_counters") {  // c = aldic_mex('frame_counters') (1 x 11)
        int c[11];
        check(aldic_frame_counters(g_ctx, c, 11), "frame_counters");
        plhs[0] = dmat(1, 11);
        for (int k = 0; k < 11; ++k) mxGetDoubles(plhs[0])[k] = c[k];
    } else if (cmd == "frame_scalars") {  // v = aldic_mex('frame_scalars')
        plhs[0] = dmat(1, 4);
        check(aldic_frame_scalars(g_ctx, mxGetDoubles(plhs[0]), 4), "frame_scalars");
    } else if (cmd == "qframe_setup") {
        // aldic_mex('qframe_setup', winsizeList, tri (1-based, nn x 3), w, tol, mu, beta,
        //           n_admm, cg_maxit, cg_rtol [, elementMinSize (0: no strain smoothing),
        //           StrainSmoothness (1e-4)])
        need(nrhs, 10, cmd.c_str());
        const std::vector<int> ws = ints(prhs[1], 0);
        const size_t nn = ws.size();
        const double *t = dbl(prhs[2], "tri"), *ww = dbl(prhs[3], "w");
        std::vector<int> tri(3 * nn);
        std::vector<double> w(3 * nn);
        for (size_t i = 0; i < nn; ++i)
            for (int k = 0; k < 3; ++k) {
                tri[3 * i + k] = (int)t[k * nn + i] - 1;
                w[3 * i + k] = ww[k * nn + i];
            }
        aldic_qframe_opts o{};
        o.tol = mxGetScalar(prhs[4]);
        o.mu = mxGetScalar(prhs[5]);
        o.beta = mxGetScalar(prhs[6]);
        o.n_admm = (int)mxGetScalar(prhs[7]);
        o.cg_maxit = (int)mxGetScalar(prhs[8]);
        o.cg_rtol = mxGetScalar(prhs[9]);
        o.smooth_h = nrhs > 10 ? mxGetScalar(prhs[10]) : 0.0;
        o.smoothness = nrhs > 11 ? mxGetScalar(prhs[11]) : 1e-4;
        check(aldic_qframe_setup(g_ctx, &o, ws.data(), tri.data(), w.data()), "qframe_setup");
    } else if (cmd == "qframe_run") {  // aldic_mex('qframe_run', U0)
        need(nrhs, 2, cmd.c_str());
        check(aldic_qframe_run(g_ctx, dbl(prhs[1], "U0")), "qframe_run");
    } else if (cmd == "qframe_get") {  // [U, F, it6, it2, times] = aldic_mex('qframe_get', nn)
        need(nrhs, 2, cmd.c_str());
        const size_t nn = (size_t)mxGetScalar(prhs[1]);
        plhs[0] = dmat(2 * nn, 1); plhs[1] = dmat(4 * nn, 1); plhs[2] = imat(nn, 1);
        plhs[3] = imat(nn, 1); plhs[4] = dmat(1, 6);
        check(aldic_qframe_get(g_ctx, mxGetDoubles(plhs[0]), mxGetDoubles(plhs[1]), iptr(plhs[2]),
                               iptr(plhs[3]), mxGetDoubles(plhs[4])),
              "qframe_get");
    } else {
        mexErrMsgIdAndTxt("aldic:cmd", "unknown command '%s'", cbuf);
    }
}
