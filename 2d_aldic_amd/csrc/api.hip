// C-ABI entry points (include/aldic.h): argument checks, host<->device staging,
// launch bucketing, mesh set-up and the device-resident frame driver.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "aldic_internal.h"

using namespace aldic;

namespace {

int fail(aldic_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

int hipfail(aldic_ctx *c, hipError_t e, const char *what)
{
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    return fail(c, ALDIC_ERR_HIP, buf);
}

#define CK(x)                                              \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return hipfail(ctx, e_, #x); \
    } while (0)

size_t dtype_size(int dt)
{
    switch (dt) {
    case ALDIC_U8: return 1;
    case ALDIC_U16: return 2;
    case ALDIC_F32: return 4;
    case ALDIC_F64: return 8;
    default: return 0;
    }
}

template <class T>
hipError_t up(DBuf &b, const T *h, size_t n, hipStream_t st)
{
    hipError_t e = b.ensure(sizeof(T) * n);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(b.p, h, sizeof(T) * n, hipMemcpyHostToDevice, st);
}

template <class T>
hipError_t down(T *h, const void *d, size_t n, hipStream_t st)
{
    if (!h) return hipSuccess;
    return hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, st);
}

// main_ALDIC.m:285-301 (FE branch): z-scores of both error norms (std with N - 1), the
// first minimum of their sum, and the vertex of the parabola through the three
// (log10 beta, sum) points around it; a minimum at an end of the list keeps that entry
// (where the reference's fit raises and its catch takes the list value).
double beta_choice(const double bl[6], const double e1[6], const double e2[6])
{
    double m1 = 0, m2 = 0;
    for (int k = 0; k < 6; ++k) { m1 += e1[k] / 6; m2 += e2[k] / 6; }
    double v1 = 0, v2 = 0;
    for (int k = 0; k < 6; ++k) {
        v1 += (e1[k] - m1) * (e1[k] - m1) / 5;
        v2 += (e2[k] - m2) * (e2[k] - m2) / 5;
    }
    double es[6];
    for (int k = 0; k < 6; ++k) es[k] = (e1[k] - m1) / std::sqrt(v1) + (e2[k] - m2) / std::sqrt(v2);
    int i = 0;
    for (int k = 1; k < 6; ++k)
        if (es[k] < es[i]) i = k;
    if (i == 0 || i == 5) return bl[i];
    const double x0 = std::log10(bl[i - 1]), x1 = std::log10(bl[i]), x2 = std::log10(bl[i + 1]);
    const double d01 = (es[i] - es[i - 1]) / (x1 - x0), d12 = (es[i + 1] - es[i]) / (x2 - x1);
    const double p1 = (d12 - d01) / (x2 - x0), p2 = d01 - p1 * (x0 + x1);
    return std::pow(10.0, -p2 / (2.0 * p1));
}

void clip_roi(const int *roi, int nx, int ny, int out[4])
{
    out[0] = roi ? std::max(1, roi[0]) : 1;
    out[1] = roi ? std::min(nx, roi[1]) : nx;
    out[2] = roi ? std::max(1, roi[2]) : 1;
    out[3] = roi ? std::min(ny, roi[3]) : ny;
}

// Normalise + gradient from the staged raw pair (device resident).
hipError_t run_prep(aldic_ctx *c)
{
    const size_t np = (size_t)c->nx * c->ny;
    hipError_t e;
    if ((e = c->f.ensure(4 * np)) || (e = c->g.ensure(4 * np)) || (e = c->gx.ensure(4 * np)) ||
        (e = c->gy.ensure(4 * np)) ||
        (e = c->stats.ensure((4 * kStatParts + 4) * sizeof(double))))
        return e;
    double *acc = c->stats.as<double>();  // partials of f, then of g
    if ((e = launch_stats(c->rawf.p, c->rawg.p, c->raw_dtype, c->nx, c->ny, c->roi, acc,
                          c->stream)))
        return e;
    const double nroi = (double)(c->roi[1] - c->roi[0] + 1) * (c->roi[3] - c->roi[2] + 1);
    return launch_prep(c->rawf.p, c->rawg.p, c->raw_dtype, c->nx, c->ny,
                       c->have_mask ? c->mask.as<uint8_t>() : nullptr, acc, nroi,
                       acc + 4 * kStatParts, c->f.as<float>(), c->g.as<float>(),
                       c->gx.as<float>(), c->gy.as<float>(), c->stream);
}

// Whether the search that will run fits its limits: the int8 searches on staged 8- or
// 16-bit pixels have their own geometry limits, the FP32 search needs its window in LDS.
bool search_fits(const aldic_ctx *c, bool raw, int ws, int sx, int sy)
{
    if (raw && c->raw_dtype == ALDIC_U8 &&
        (search_i8_supported(ws, sx, sy) || search_mt_supported(ws, sx, sy, 8)))
        return true;
    if (raw && c->raw_dtype == ALDIC_U16 && search_mt_supported(ws, sx, sy, 16)) return true;
    return search_lds_bytes(ws, sx, sy) <= 160 * 1024 - 256;
}

// 8- or 16-bit raw images behind f, g: the search runs on them exactly (int8 matrix cores)
void search_raw8(aldic_ctx *c, SearchArgs &a)
{
    if (!c->images_from_raw) return;
    if (c->raw_dtype == ALDIC_U8) {
        a.f8 = c->rawf.as<uint8_t>();
        a.g8 = c->rawg.as<uint8_t>();
    } else if (c->raw_dtype == ALDIC_U16) {
        a.f16 = c->rawf.as<uint16_t>();
        a.g16 = c->rawg.as<uint16_t>();
    } else {
        return;
    }
    a.ms = c->stats.as<double>() + 4 * kStatParts;
    a.planes = &c->boxp;
}

int stage_raw(aldic_ctx *ctx, const void *f, const void *g, int nx, int ny, int dtype,
              const uint8_t *mask, const int *roi)
{
    if (!f || !g || nx < 8 || ny < 8) return fail(ctx, ALDIC_ERR_ARG, "images: bad arguments");
    if ((size_t)nx * ny > kMaxPixels || nx >= (1 << 22))
        return fail(ctx, ALDIC_ERR_ARG, "images: more than 2^29 pixels or 2^22 columns");
    const size_t es = dtype_size(dtype);
    if (!es) return fail(ctx, ALDIC_ERR_ARG, "images: unknown dtype");
    const size_t np = (size_t)nx * ny;
    ctx->have_images = false;  // the normalised planes are made from the new pair later
    // 64 spare bytes: the 8-bit search reads whole aligned dwords around each row segment
    CK(ctx->rawf.ensure(es * np + 64));
    CK(ctx->rawg.ensure(es * np + 64));
    CK(hipMemcpyAsync(ctx->rawf.p, f, es * np, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(ctx->rawg.p, g, es * np, hipMemcpyHostToDevice, ctx->stream));
    ctx->have_mask = mask != nullptr;
    if (mask) CK(up(ctx->mask, mask, np, ctx->stream));
    ctx->nx = nx;
    ctx->ny = ny;
    ctx->raw_dtype = dtype;
    clip_roi(roi, nx, ny, ctx->roi);
    if (ctx->roi[1] - ctx->roi[0] < 1 || ctx->roi[3] - ctx->roi[2] < 1)
        return fail(ctx, ALDIC_ERR_ARG, "images: empty ROI");
    return ALDIC_OK;
}

// Pixels-per-lane class of each subset; launches each class once.
int icgn_dispatch(aldic_ctx *ctx, bool six, IcgnArgs a, const int *ws_host)
{
    const int S = a.S;
    std::vector<int> cls(S);
    for (int j = 0; j < S; ++j) {
        const int w = ws_host ? ws_host[j] : a.ws_all;
        cls[j] = icgn_ppl_for(w);
        if (w < 2 || cls[j] == 0)
            return fail(ctx, ALDIC_ERR_ARG, "icgn: subset size must be 2..64");
    }
    const int classes[4] = {5, 8, 18, 72};
    std::vector<int> all(S);
    DBuf &lst = ctx->scratch[7];
    bool single = std::all_of(cls.begin(), cls.end(), [&](int v) { return v == cls[0]; });
    int *iters = a.iters;
    std::vector<int> pending;
    if (single) {
        a.list = nullptr;
        a.nlist = S;
        CK(six ? launch_icgn6(a, cls[0], ctx->stream) : launch_icgn2(a, cls[0], ctx->stream));
    } else {
        for (int ci = 0; ci < 4; ++ci) {
            std::vector<int> l;
            for (int j = 0; j < S; ++j)
                if (cls[j] == classes[ci]) l.push_back(j);
            if (l.empty()) continue;
            CK(up(lst, l.data(), l.size(), ctx->stream));
            a.list = lst.as<int>();
            a.nlist = (int)l.size();
            CK(six ? launch_icgn6(a, classes[ci], ctx->stream)
                   : launch_icgn2(a, classes[ci], ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));  // list buffer is reused
        }
    }
    // windows the mask rule enlarged to round(sqrt(2) w) beyond their class (status -2,
    // funICGN.m:57-65): one relaunch in the class that holds the enlarged window; when no
    // class does (enlarged window above 67 x 67 px) the subset keeps status -2
    std::vector<int> it(S);
    CK(down(it.data(), iters, S, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    for (int ci = 0; ci < 4; ++ci) {
        std::vector<int> l2;
        for (int j = 0; j < S; ++j) {
            if (it[j] != -2) continue;
            const int w = ws_host ? ws_host[j] : a.ws_all;
            const int big = icgn_ppl_for((int)std::floor(std::sqrt(2.0) * w + 0.5));
            if (big == classes[ci] && big != cls[j]) l2.push_back(j);
        }
        if (l2.empty()) continue;
        CK(up(lst, l2.data(), l2.size(), ctx->stream));
        a.list = lst.as<int>();
        a.nlist = (int)l2.size();
        CK(six ? launch_icgn6(a, classes[ci], ctx->stream)
               : launch_icgn2(a, classes[ci], ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
    }
    return ALDIC_OK;
}

}  // namespace

extern "C" {

int aldic_create(aldic_ctx **out, int device)
{
    if (!out) return ALDIC_ERR_ARG;
    *out = nullptr;
    aldic_ctx *ctx = new (std::nothrow) aldic_ctx();
    if (!ctx) return ALDIC_ERR_ARG;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = noalloc_init(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    for (int k = 0; k < 10 && e == hipSuccess; ++k) e = hipEventCreate(&ctx->ev[k]);
    if (e != hipSuccess) {
        delete ctx;
        return ALDIC_ERR_HIP;
    }
    *out = ctx;
    return ALDIC_OK;
}

void aldic_destroy(aldic_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DBuf *bufs[] = {&ctx->f, &ctx->g, &ctx->gx, &ctx->gy, &ctx->rawf, &ctx->rawg, &ctx->mask,
                    &ctx->stats, &ctx->mesh.coords, &ctx->mesh.elems, &ctx->mesh.n2e_ptr,
                    &ctx->mesh.n2e, &ctx->mesh.rowptr, &ctx->mesh.col, &ctx->mesh.aval,
                    &ctx->mesh.mval, &ctx->mesh.diag, &ctx->mesh.gvec, &ctx->mesh.fdV, &ctx->mesh.fdlam};
    for (DBuf *b : bufs) b->release();
    for (auto &b : ctx->scratch) b.release();
    for (auto &b : ctx->lic) b.release();
    FrameState &F = ctx->fr;
    DBuf *fb[] = {&F.coords, &F.U0, &F.U6, &F.F6, &F.H21, &F.it6, &F.zncc, &F.Ucur, &F.Fcur,
                  &F.U1, &F.F1, &F.U2, &F.F2, &F.udual, &F.vdual, &F.it2, &F.bad, &F.tmp,
                  &F.part, &F.cc, &F.rl, &F.F2b};
    for (DBuf *b : fb) b->release();
    if (F.ev_skern) (void)hipEventDestroy(F.ev_skern);
    if (F.ev_ikern) (void)hipEventDestroy(F.ev_ikern);
    for (aldic::GridfitDev *g : {&F.gf, &ctx->gf}) {
        g->w.release();
        g->a9.release();
        g->dinv.release();
    }
    F.gfstat.release();
    F.nat_err.release();
    F.nbad.release();
    outlier_work_free(ctx->outl);
    ctx->outl = nullptr;
    cnn_work_free(ctx->cnn);
    ctx->cnn = nullptr;
    mg_work_free(ctx->mg);
    ctx->mg = nullptr;
    for (DBuf &bb : ctx->smq) bb.release();
    F.seed.release();
    for (DBuf *bb : {&F.P6, &F.F0, &F.lflag, &ctx->qf.sgidx, &ctx->qf.sw[0], &ctx->qf.sw[1]}) bb->release();
    if (F.nfb_h) (void)hipHostFree(F.nfb_h);
    ctx->qf.nat_err.release();
    if (F.nat_err_h) (void)hipHostFree(F.nat_err_h);
    if (ctx->qf.nat_err_h) (void)hipHostFree(ctx->qf.nat_err_h);
    for (aldic::NatGrid *g : {&F.nat, &ctx->nat, &ctx->quad.nat})
        for (DBuf *bb : {&g->bstart, &g->bnode, &g->rstart, &g->rnode, &g->work}) bb->release();
    if (F.gfh) (void)hipHostFree(F.gfh);
    if (F.hostd) (void)hipHostFree(F.hostd);
    for (DBuf *bb : {&F.Us, &F.Fs, &F.U1s, &F.U1old, &F.U2old, &F.errs}) bb->release();
    aldic::QuadDev &Q = ctx->quad;
    for (DBuf *bb : {&Q.m.coords, &Q.m.n2e_ptr, &Q.m.rowptr, &Q.m.col, &Q.m.aval, &Q.m.mval,
                     &Q.m.diag, &Q.m.gvec, &Q.elems, &Q.n2e, &Q.code, &Q.cval, &ctx->boxp})
        bb->release();
    QFrameState &QF = ctx->qf;
    for (DBuf *bb : {&QF.ws, &QF.U0, &QF.U6, &QF.F6, &QF.H21, &QF.it6, &QF.U1, &QF.F1, &QF.U2,
                     &QF.F2a, &QF.F2b, &QF.udual, &QF.vdual, &QF.it2, &QF.bad, &QF.part,
                     &QF.known, &QF.tri, &QF.w, &QF.Fgp, &QF.lists, &QF.rl, &QF.b, &QF.cc,
                     &QF.pstat})
        bb->release();
    if (QF.hstat) (void)hipHostFree(QF.hstat);
    for (auto &e : QF.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : F.ev) (void)hipEventDestroy(e);
    F.ev.clear();
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *aldic_last_error(const aldic_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int aldic_set_images(aldic_ctx *ctx, const void *f, const void *g, int nx, int ny, int dtype,
                     const uint8_t *mask, const int *roi)
{
    if (!ctx) return ALDIC_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    int r = stage_raw(ctx, f, g, nx, ny, dtype, mask, roi);
    if (r) return r;
    CK(run_prep(ctx));
    CK(hipStreamSynchronize(ctx->stream));
    ctx->have_images = true;
    ctx->images_from_raw = true;
    return ALDIC_OK;
}

int aldic_set_normalized(aldic_ctx *ctx, const double *f, const double *g, int nx, int ny)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!f || !g || nx < 8 || ny < 8) return fail(ctx, ALDIC_ERR_ARG, "set_normalized: bad arguments");
    if ((size_t)nx * ny > kMaxPixels || nx >= (1 << 22))
        return fail(ctx, ALDIC_ERR_ARG, "set_normalized: more than 2^29 pixels or 2^22 columns");
    CK(hipSetDevice(ctx->device));
    const size_t np = (size_t)nx * ny;
    std::vector<float> tmp(np);
    CK(ctx->f.ensure(4 * np));
    CK(ctx->g.ensure(4 * np));
    CK(ctx->gx.ensure(4 * np));
    CK(ctx->gy.ensure(4 * np));
    for (size_t i = 0; i < np; ++i) tmp[i] = (float)f[i];
    CK(hipMemcpy(ctx->f.p, tmp.data(), 4 * np, hipMemcpyHostToDevice));
    for (size_t i = 0; i < np; ++i) tmp[i] = (float)g[i];
    CK(hipMemcpy(ctx->g.p, tmp.data(), 4 * np, hipMemcpyHostToDevice));
    ctx->nx = nx;
    ctx->ny = ny;
    CK(launch_gradient(ctx->f.as<float>(), nx, ny, ctx->gx.as<float>(), ctx->gy.as<float>(),
                       ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    ctx->have_images = true;
    ctx->images_from_raw = false;
    return ALDIC_OK;
}

int aldic_set_gradients(aldic_ctx *ctx, const double *gx, const double *gy)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "set_gradients: no images");
    if (!gx || !gy) return fail(ctx, ALDIC_ERR_ARG, "set_gradients: NULL");
    CK(hipSetDevice(ctx->device));
    const size_t np = (size_t)ctx->nx * ctx->ny;
    std::vector<float> tmp(np);
    const int nx = ctx->nx, ny = ctx->ny;
    const double *src[2] = {gx, gy};
    DBuf *dst[2] = {&ctx->gx, &ctx->gy};
    for (int k = 0; k < 2; ++k) {
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const bool in = x >= 3 && x <= nx - 4 && y >= 3 && y <= ny - 4;
                tmp[(size_t)y * nx + x] = in ? (float)src[k][(size_t)y * nx + x] : 0.f;
            }
        CK(hipMemcpy(dst[k]->p, tmp.data(), 4 * np, hipMemcpyHostToDevice));
    }
    return ALDIC_OK;
}

int aldic_get_images(aldic_ctx *ctx, double *f, double *g, double *gx, double *gy)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "get_images: no images");
    CK(hipSetDevice(ctx->device));
    const size_t np = (size_t)ctx->nx * ctx->ny;
    std::vector<float> tmp(np);
    double *dst[4] = {f, g, gx, gy};
    DBuf *src[4] = {&ctx->f, &ctx->g, &ctx->gx, &ctx->gy};
    for (int k = 0; k < 4; ++k) {
        if (!dst[k]) continue;
        CK(hipMemcpy(tmp.data(), src[k]->p, 4 * np, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < np; ++i) dst[k][i] = tmp[i];
    }
    return ALDIC_OK;
}

int aldic_search_grid(int nx, int ny, const int *grid, int ws, int step, const int *s, int *out)
{
    if (!grid || !s || !out || ws < 2 || step < 1 || s[0] < 0 || s[1] < 0) return ALDIC_ERR_ARG;
    int gx1 = grid[0], gx2 = grid[1], gy1 = grid[2], gy2 = grid[3];
    // funIntegerSearch.m:14-25
    while (gx1 - s[0] - 0.5 * ws < 1) ++gx1;
    while (gy1 - s[1] - 0.5 * ws < 1) ++gy1;
    while (gx2 + 1.5 * ws + s[0] > nx) --gx2;
    while (gy2 + 1.5 * ws + s[1] > ny) --gy2;
    // :90-91 XList = gridx(1) : step : gridx(2) - step
    const int Nx = gx2 - step >= gx1 ? (gx2 - step - gx1) / step + 1 : 0;
    const int Ny = gy2 - step >= gy1 ? (gy2 - step - gy1) / step + 1 : 0;
    out[0] = gx1;
    out[1] = gy1;
    out[2] = Nx;
    out[3] = Ny;
    return ALDIC_OK;
}

int aldic_integer_search(aldic_ctx *ctx, const int *grid, int ws, int step, const int *s,
                         double *x, double *y, double *u, double *v, double *ccmax, double *qf,
                         int *status)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "integer_search: no images");
    int gr[4];
    if (aldic_search_grid(ctx->nx, ctx->ny, grid, ws, step, s, gr))
        return fail(ctx, ALDIC_ERR_ARG, "integer_search: bad grid arguments");
    const int S = gr[2] * gr[3];
    if (S <= 0) return ALDIC_OK;
    if (!search_fits(ctx, ctx->images_from_raw, ws, s[0], s[1]))
        return fail(ctx, ALDIC_ERR_ARG, "integer_search: window too large for LDS");
    CK(hipSetDevice(ctx->device));
    DBuf *sc = ctx->scratch;
    for (int k = 0; k < 6; ++k) CK(sc[k].ensure(sizeof(double) * 2 * (size_t)S));
    SearchArgs a{};
    a.f = ctx->f.as<float>();
    a.g = ctx->g.as<float>();
    a.nx = ctx->nx;
    a.ny = ctx->ny;
    a.x1 = gr[0];
    a.y1 = gr[1];
    a.Nx = gr[2];
    a.Ny = gr[3];
    a.ws = ws;
    a.step = step;
    a.sx = s[0];
    a.sy = s[1];
    a.x = sc[0].as<double>();
    a.y = sc[1].as<double>();
    a.u = sc[2].as<double>();
    a.v = sc[3].as<double>();
    a.ccmax = sc[4].as<double>();
    a.qf = qf ? sc[5].as<double>() : nullptr;
    CK(ctx->scratch[6].ensure(sizeof(int) * (size_t)S));
    a.status = ctx->scratch[6].as<int>();
    a.U0 = nullptr;
    search_raw8(ctx, a);
    if (qf && qfactor_lds_bytes(ws, s[0], s[1]) > 150 * 1024)
        return fail(ctx, ALDIC_ERR_ARG, "integer_search: full NCC map too large for q-factors");
    CK(launch_search(a, ctx->stream));
    if (qf) CK(launch_qfactors(a, ctx->stream));
    CK(down(x, a.x, S, ctx->stream));
    CK(down(y, a.y, S, ctx->stream));
    CK(down(u, a.u, S, ctx->stream));
    CK(down(v, a.v, S, ctx->stream));
    CK(down(ccmax, a.ccmax, S, ctx->stream));
    if (qf) CK(down(qf, a.qf, 2 * (size_t)S, ctx->stream));
    CK(down(status, a.status, S, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_smooth_strain_quadtree(aldic_ctx *ctx, int nn, const double *coords, double h, double smoothness,
                                 double filter_std, int filter_size, double *F, int *iters)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (nn < 1 || !coords || !F || !(h > 0) || !(smoothness > 0) || filter_std < 0 || filter_size < 0)
        return fail(ctx, ALDIC_ERR_ARG, "smooth_strain_quadtree: bad arguments");
    // funSmoothStrainQuadtree.m:40-61: std 0 -> 0.5, size 0 -> 2 ceil(2 std) + 1
    const double sd = filter_std == 0 ? 0.5 : filter_std;
    const int fs = filter_size == 0 ? 2 * (int)std::ceil(2.0 * sd) + 1 : filter_size;
    if (fs != 3) return fail(ctx, ALDIC_ERR_ARG, "smooth_strain_quadtree: only the 3 x 3 filter is built");
    // the grid min:h:max of each axis; every node must sit on it (the reference's find)
    int nx = 0, ny = 0;
    std::vector<int> gidx;
    if (!smooth_grid(nn, coords, h, nx, ny, gidx))
        return fail(ctx, ALDIC_ERR_ARG, "smooth_strain_quadtree: a node is off the grid min:h:max, or the grid "
                                        "has fewer than 3 points per axis or more than 2^26");
    CK(hipSetDevice(ctx->device));
    DBuf *w = ctx->smq;
    CK(w[2].ensure(sizeof(int) * (size_t)nn));
    CK(w[3].ensure(sizeof(double) * 4 * (size_t)nn));
    CK(hipMemcpyAsync(w[2].p, gidx.data(), sizeof(int) * (size_t)nn, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(w[3].p, F, sizeof(double) * 4 * (size_t)nn, hipMemcpyHostToDevice, ctx->stream));
    int it[2] = {0, 0};
    const hipError_t e = smooth_strain_qt(nn, w[2].as<int>(), nx, ny, smoothness, sd, 1e-12, 20000,
                                          w[3].as<double>(), w, it, ctx->stream);
    if (e == hipErrorUnknown) return fail(ctx, ALDIC_ERR_NUMERIC, "smooth_strain_quadtree: non-finite residual");
    CK(e);
    CK(down(F, w[3].as<double>(), 4 * (size_t)nn, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (iters) { iters[0] = it[0]; iters[1] = it[1]; }
    return ALDIC_OK;
}

int aldic_integer_search_mg(aldic_ctx *ctx, const int *grid, int ws, int step, long cap, int *dims, double *x,
                            double *y, double *u, double *v, double *ccmax, long temp_cap, double *temp)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!grid || !dims || ws < 1 || step < 1 || cap < 0 || (cap > 0 && (!x || !y || !u || !v || !ccmax)))
        return fail(ctx, ALDIC_ERR_ARG, "integer_search_mg: bad arguments");
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "integer_search_mg: no images");
    CK(hipSetDevice(ctx->device));
    if (!ctx->mg) ctx->mg = mg_work_new();
    if (!ctx->mg) return fail(ctx, ALDIC_ERR_HIP, "integer_search_mg: out of host memory");
    MgResult r;
    std::string msg;
    const int rc = integer_search_mg(*ctx->mg, ctx->f.as<float>(), ctx->g.as<float>(), ctx->nx, ctx->ny, grid, ws,
                                     step, r, msg, ctx->stream);
    if (rc == 1) return fail(ctx, ALDIC_ERR_ARG, msg.c_str());
    if (rc) return fail(ctx, ALDIC_ERR_HIP, msg.c_str());
    const int Nx = (int)r.xlist.size(), Ny = (int)r.ylist.size(), nr = (int)r.tempx.size();
    int nnan = 0;
    for (int q = 0; q < nr; ++q) nnan += std::isnan(r.tempu[q]) ? 1 : 0;
    dims[0] = Nx;
    dims[1] = Ny;
    dims[2] = r.levels;
    dims[3] = nr;
    dims[4] = nnan;
    dims[5] = r.fit_iters[0];
    if ((long)Nx * Ny > cap) return fail(ctx, ALDIC_ERR_ARG, "integer_search_mg: output capacity too small");
    if (temp && 5L * nr > temp_cap)
        return fail(ctx, ALDIC_ERR_ARG, "integer_search_mg: region output capacity too small");
    for (int j = 0; j < Ny; ++j)
        for (int i = 0; i < Nx; ++i) {
            const long k = (long)j * Nx + i;
            x[k] = r.xlist[i];
            y[k] = r.ylist[j];
            u[k] = r.u[k];
            v[k] = r.v[k];
            ccmax[k] = r.cc[k];
        }
    if (temp)
        for (int q = 0; q < nr; ++q) {
            temp[q] = r.tempx[q];
            temp[nr + q] = r.tempy[q];
            temp[2L * nr + q] = r.tempu[q];
            temp[3L * nr + q] = r.tempv[q];
            temp[4L * nr + q] = r.tempphi[q];
        }
    return ALDIC_OK;
}

static int icgn6_impl(aldic_ctx *ctx, int S, const double *coords, const double *U0,
                      const double *F0, const int *ws, int ws_all, double tol, double *U, double *F,
                      double *H21, int *iters, double *zncc)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "icgn6: no images");
    if (S < 0 || (S > 0 && (!coords || !U0 || !U || !F || !H21 || !iters)))
        return fail(ctx, ALDIC_ERR_ARG, "icgn6: bad arguments");
    if (S == 0) return ALDIC_OK;
    CK(hipSetDevice(ctx->device));
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], coords, 2 * (size_t)S, ctx->stream));
    CK(up(sc[1], U0, 2 * (size_t)S, ctx->stream));
    if (F0) CK(up(sc[8], F0, 4 * (size_t)S, ctx->stream));
    CK(sc[2].ensure(sizeof(double) * 2 * S));
    CK(sc[3].ensure(sizeof(double) * 4 * S));
    CK(sc[4].ensure(sizeof(double) * 21 * (size_t)S));
    CK(sc[5].ensure(sizeof(double) * (size_t)S + sizeof(int) * (size_t)S));
    if (ws) CK(up(sc[6], ws, S, ctx->stream));
    IcgnArgs a{};
    a.f = ctx->f.as<float>();
    a.g = ctx->g.as<float>();
    a.gx = ctx->gx.as<float>();
    a.gy = ctx->gy.as<float>();
    a.nx = ctx->nx;
    a.ny = ctx->ny;
    a.S = S;
    a.coords = sc[0].as<double>();
    a.U0 = sc[1].as<double>();
    a.F0 = F0 ? sc[8].as<double>() : nullptr;
    a.ws = ws ? sc[6].as<int>() : nullptr;
    a.ws_all = ws_all;
    a.tol = tol;
    a.U = sc[2].as<double>();
    a.F = sc[3].as<double>();
    a.H21 = sc[4].as<double>();
    a.zncc = sc[5].as<double>();
    a.iters = reinterpret_cast<int *>(sc[5].as<double>() + S);
    int r = icgn_dispatch(ctx, true, a, ws);
    if (r) return r;
    CK(down(U, a.U, 2 * (size_t)S, ctx->stream));
    CK(down(F, a.F, 4 * (size_t)S, ctx->stream));
    CK(down(H21, a.H21, 21 * (size_t)S, ctx->stream));
    CK(down(iters, a.iters, S, ctx->stream));
    if (zncc) CK(down(zncc, a.zncc, S, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_icgn6(aldic_ctx *ctx, int S, const double *coords, const double *U0, const int *ws,
                int ws_all, double tol, double *U, double *F, double *H21, int *iters,
                double *zncc)
{
    return icgn6_impl(ctx, S, coords, U0, nullptr, ws, ws_all, tol, U, F, H21, iters, zncc);
}

int aldic_icgn6_start(aldic_ctx *ctx, int S, const double *coords, const double *P0,
                      const int *ws, int ws_all, double tol, double *U, double *F, double *H21,
                      int *iters, double *zncc)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (S < 0 || (S > 0 && !P0)) return fail(ctx, ALDIC_ERR_ARG, "icgn6_start: bad arguments");
    try {
        std::vector<double> U0(2 * (size_t)S), F0(4 * (size_t)S);
        for (int j = 0; j < S; ++j) {
            for (int q = 0; q < 4; ++q) F0[4 * (size_t)j + q] = P0[6 * (size_t)j + q];
            U0[2 * (size_t)j] = P0[6 * (size_t)j + 4];
            U0[2 * (size_t)j + 1] = P0[6 * (size_t)j + 5];
        }
        return icgn6_impl(ctx, S, coords, U0.data(), F0.data(), ws, ws_all, tol, U, F, H21, iters,
                          zncc);
    } catch (...) {
        return fail(ctx, ALDIC_ERR_HIP, "icgn6_start: host allocation failed");
    }
}

int aldic_icgn2(aldic_ctx *ctx, int S, const double *coords, const int *ws, int ws_all,
                const double *H21, double mu, const double *vdual, const double *Uold,
                const double *Fold, double tol, double *U, int *iters)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "icgn2: no images");
    if (S < 0 || (S > 0 && (!coords || !H21 || !vdual || !Uold || !Fold || !U || !iters)))
        return fail(ctx, ALDIC_ERR_ARG, "icgn2: bad arguments");
    if (S == 0) return ALDIC_OK;
    CK(hipSetDevice(ctx->device));
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], coords, 2 * (size_t)S, ctx->stream));
    CK(up(sc[1], H21, 21 * (size_t)S, ctx->stream));
    CK(up(sc[2], vdual, 2 * (size_t)S, ctx->stream));
    CK(up(sc[3], Uold, 2 * (size_t)S, ctx->stream));
    CK(up(sc[4], Fold, 4 * (size_t)S, ctx->stream));
    CK(sc[5].ensure(sizeof(double) * 2 * (size_t)S + sizeof(int) * (size_t)S));
    if (ws) CK(up(sc[6], ws, S, ctx->stream));
    IcgnArgs a{};
    a.f = ctx->f.as<float>();
    a.g = ctx->g.as<float>();
    a.gx = ctx->gx.as<float>();
    a.gy = ctx->gy.as<float>();
    a.nx = ctx->nx;
    a.ny = ctx->ny;
    a.S = S;
    a.coords = sc[0].as<double>();
    a.Hin = sc[1].as<double>();
    a.vdual = sc[2].as<double>();
    a.Uold = sc[3].as<double>();
    a.Fold = sc[4].as<double>();
    a.mu = mu;
    a.ws = ws ? sc[6].as<int>() : nullptr;
    a.ws_all = ws_all;
    a.tol = tol;
    a.U = sc[5].as<double>();
    a.iters = reinterpret_cast<int *>(sc[5].as<double>() + 2 * (size_t)S);
    int r = icgn_dispatch(ctx, false, a, ws);
    if (r) return r;
    CK(down(U, a.U, 2 * (size_t)S, ctx->stream));
    CK(down(iters, a.iters, S, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_set_mesh(aldic_ctx *ctx, const double *coords, int nn, const int *elems, int E)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (nn <= 0 || E < 0 || !coords || (E > 0 && !elems))
        return fail(ctx, ALDIC_ERR_ARG, "set_mesh: bad arguments");
    for (int e = 0; e < 4 * E; ++e)
        if (elems[e] < 0 || elems[e] >= nn) return fail(ctx, ALDIC_ERR_ARG, "set_mesh: node index");
    CK(hipSetDevice(ctx->device));
    MeshDev &m = ctx->mesh;
    // node -> element incidences
    std::vector<int> cnt(nn + 1, 0);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < 4; ++a) cnt[elems[4 * e + a] + 1]++;
    for (int i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
    std::vector<N2E> n2e(cnt[nn]);
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < 4; ++a) {
            N2E r{};
            r.e = e;
            r.a = a;
            n2e[fillp[elems[4 * e + a]]++] = r;
        }
    // scalar pattern
    std::vector<int> rowptr(nn + 1, 0), col;
    col.reserve((size_t)nn * 9);
    for (int i = 0; i < nn; ++i) {
        std::vector<int> nb;
        for (int q = cnt[i]; q < cnt[i + 1]; ++q)
            for (int b = 0; b < 4; ++b) nb.push_back(elems[4 * n2e[q].e + b]);
        if (nb.empty()) nb.push_back(i);
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
        for (int q = cnt[i]; q < cnt[i + 1]; ++q)
            for (int b = 0; b < 4; ++b) {
                const int c = elems[4 * n2e[q].e + b];
                n2e[q].slot[b] = rowptr[i] + (int)(std::lower_bound(nb.begin(), nb.end(), c) - nb.begin());
            }
        col.insert(col.end(), nb.begin(), nb.end());
        rowptr[i + 1] = (int)col.size();
    }
    m.nn = nn;
    m.E = E;
    m.nnz = (int)col.size();
    CK(up(m.coords, coords, 2 * (size_t)nn, ctx->stream));
    if (E > 0) CK(up(m.elems, elems, 4 * (size_t)E, ctx->stream));
    else CK(m.elems.ensure(16));
    CK(up(m.n2e_ptr, cnt.data(), cnt.size(), ctx->stream));
    if (!n2e.empty()) CK(up(m.n2e, n2e.data(), n2e.size(), ctx->stream));
    else CK(m.n2e.ensure(sizeof(N2E)));
    CK(up(m.rowptr, rowptr.data(), rowptr.size(), ctx->stream));
    CK(up(m.col, col.data(), col.size(), ctx->stream));
    CK(m.aval.ensure(sizeof(double) * col.size()));
    CK(m.mval.ensure(sizeof(double) * col.size()));
    CK(m.diag.ensure(sizeof(double) * nn));
    CK(m.gvec.ensure(sizeof(double) * 2 * nn));
    CK(hipStreamSynchronize(ctx->stream));
    m.beta = m.mu = m.alpha = -1;
    // structured lattice? (x-fastest uniform grid with the standard Q4 connectivity)
    m.structured = false;
    {
        int gx = 1;
        while (gx < nn && coords[2 * gx + 1] == coords[1] && coords[2 * gx] > coords[2 * (gx - 1)]) ++gx;
        const int gy = gx > 0 ? nn / gx : 0;
        bool ok = gx >= 2 && gy >= 2 && gx * gy == nn && E == (gx - 1) * (gy - 1);
        const double hx = ok ? coords[2] - coords[0] : 0, hy = ok ? coords[2 * gx + 1] - coords[1] : 0;
        ok = ok && hx > 0 && hy > 0;
        for (int i = 0; ok && i < nn; ++i) {
            const double ex = coords[0] + hx * (i % gx), ey = coords[1] + hy * (i / gx);
            ok = std::fabs(coords[2 * i] - ex) <= 1e-9 * std::fabs(hx) * gx &&
                 std::fabs(coords[2 * i + 1] - ey) <= 1e-9 * std::fabs(hy) * gy;
        }
        for (int e = 0; ok && e < E; ++e) {
            const int l = e / (gx - 1), k = e % (gx - 1), n0 = l * gx + k;
            ok = elems[4 * e] == n0 && elems[4 * e + 1] == n0 + 1 &&
                 elems[4 * e + 2] == n0 + gx + 1 && elems[4 * e + 3] == n0 + gx;
        }
        if (ok) {
            // closed-form generalised eigenbases of the 1-D FE matrices (DCT-I vectors,
            // M1-orthonormal): V[i][k] = cos(pi k i / (n-1)) / sqrt(m_k w_k)
            // fdV = [Vx, Vy, Vx^T, Vy^T] (row-major)
            std::vector<double> V(2 * ((size_t)gx * gx + (size_t)gy * gy)), lam(gx + gy);
            auto basis = [](int n, double h, double *Vo, double *lo) {
                const double pi = 3.14159265358979323846;
                for (int k = 0; k < n; ++k) {
                    const double th = pi * k / (n - 1), c = std::cos(th);
                    const double mk = h / 6.0 * (4.0 + 2.0 * c);
                    const double wk = (k == 0 || k == n - 1) ? (n - 1.0) : 0.5 * (n - 1.0);
                    const double sc = 1.0 / std::sqrt(mk * wk);
                    lo[k] = (2.0 / h) * (1.0 - c) / mk;
                    for (int i = 0; i < n; ++i) Vo[(size_t)i * n + k] = std::cos(th * i) * sc;
                }
            };
            basis(gx, hx, V.data(), lam.data());
            basis(gy, hy, V.data() + (size_t)gx * gx, lam.data() + gx);
            const size_t oT = (size_t)gx * gx + (size_t)gy * gy;
            for (int i = 0; i < gx; ++i)
                for (int k = 0; k < gx; ++k) V[oT + (size_t)k * gx + i] = V[(size_t)i * gx + k];
            for (int i = 0; i < gy; ++i)
                for (int k = 0; k < gy; ++k)
                    V[oT + (size_t)gx * gx + (size_t)k * gy + i] = V[(size_t)gx * gx + (size_t)i * gy + k];
            CK(up(m.fdV, V.data(), V.size(), ctx->stream));
            CK(up(m.fdlam, lam.data(), lam.size(), ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));
            m.structured = true;
            m.gx = gx;
            m.gy = gy;
            m.hx = hx;
            m.hy = hy;
        }
    }
    ctx->have_mesh = true;
    return ALDIC_OK;
}

int aldic_subpb2(aldic_ctx *ctx, double beta, double mu, double alpha, const double *U,
                 const double *F, const double *udual, const double *vdual, double rtol,
                 int maxit, double *Uhat, double *stats)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_mesh) return fail(ctx, ALDIC_ERR_STATE, "subpb2: no mesh");
    if (!U || !F || !udual || !vdual || !Uhat || maxit < 1 || !(rtol > 0))
        return fail(ctx, ALDIC_ERR_ARG, "subpb2: bad arguments");
    CK(hipSetDevice(ctx->device));
    MeshDev &m = ctx->mesh;
    const size_t nn = m.nn;
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], U, 2 * nn, ctx->stream));
    CK(up(sc[1], F, 4 * nn, ctx->stream));
    CK(up(sc[2], udual, 4 * nn, ctx->stream));
    CK(up(sc[3], vdual, 2 * nn, ctx->stream));
    CK(sc[4].ensure(sizeof(double) * 2 * nn));
    CK(sc[5].ensure(sizeof(double) * 2 * nn));
    if (m.beta != beta || m.mu != mu || m.alpha != alpha) CK(fem_assemble(m, beta, mu, alpha, ctx->stream));
    CK(fem_rhs(m, beta, mu, sc[0].as<double>(), sc[1].as<double>(), sc[2].as<double>(),
               sc[3].as<double>(), sc[4].as<double>(), ctx->stream));
    static thread_local DBuf work[8];
    double st[2] = {0, 0};
    CK(fem_solve(m, sc[4].as<double>(), sc[5].as<double>(), rtol, maxit, work, st, ctx->stream, 8));
    CK(down(Uhat, sc[5].p, 2 * nn, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (stats) { stats[0] = st[0]; stats[1] = st[1]; }
    return ALDIC_OK;
}

int aldic_set_quadtree(aldic_ctx *ctx, const double *coords, int nn, const int *elems, int E,
                       const int *dirichlet, int nd)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (nn <= 0 || E < 0 || nd < 0 || !coords || (E > 0 && !elems) || (nd > 0 && !dirichlet))
        return fail(ctx, ALDIC_ERR_ARG, "set_quadtree: bad arguments");
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < 8; ++a) {
            const int v = elems[8 * e + a];
            if (v >= nn || (v < 0 && (a < 4 || v != -1)))
                return fail(ctx, ALDIC_ERR_ARG, "set_quadtree: node index");
        }
    for (int k = 0; k < nd; ++k)
        if (dirichlet[k] < 0 || dirichlet[k] >= nn)
            return fail(ctx, ALDIC_ERR_ARG, "set_quadtree: dirichlet node index");
    CK(hipSetDevice(ctx->device));
    aldic::QuadDev &q = ctx->quad;
    aldic::MeshDev &m = q.m;
    std::vector<int> cnt(nn + 1, 0);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < 8; ++a)
            if (elems[8 * e + a] >= 0) cnt[elems[8 * e + a] + 1]++;
    for (int i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
    std::vector<aldic::N2E8> n2e(cnt[nn]);
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < 8; ++a) {
            const int v = elems[8 * e + a];
            if (v < 0) continue;
            aldic::N2E8 r{};
            r.e = e;
            r.a = a;
            n2e[fillp[v]++] = r;
        }
    std::vector<int> rowptr(nn + 1, 0), col;
    col.reserve((size_t)nn * 13);
    for (int i = 0; i < nn; ++i) {
        std::vector<int> nb{i};
        for (int p = cnt[i]; p < cnt[i + 1]; ++p)
            for (int b = 0; b < 8; ++b)
                if (elems[8 * n2e[p].e + b] >= 0) nb.push_back(elems[8 * n2e[p].e + b]);
        std::sort(nb.begin(), nb.end());
        nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
        for (int p = cnt[i]; p < cnt[i + 1]; ++p)
            for (int b = 0; b < 8; ++b) {
                const int c = elems[8 * n2e[p].e + b];
                n2e[p].slot[b] = c < 0 ? -1
                                       : rowptr[i] + (int)(std::lower_bound(nb.begin(), nb.end(), c) - nb.begin());
            }
        col.insert(col.end(), nb.begin(), nb.end());
        rowptr[i + 1] = (int)col.size();
    }
    // roles (Subpb2Quadtree.m:165-172): the free DOFs are the nodes of the elements
    // minus the first entry of their sorted list when no element lacks a midside node
    // (the list is built assuming a 0 there), minus the Dirichlet nodes
    std::vector<int> code(nn, 2);
    bool absent = false;
    int first = nn;
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < 8; ++a) {
            const int v = elems[8 * e + a];
            if (v < 0) absent = true;
            else { code[v] = 0; first = std::min(first, v); }
        }
    if (!absent && first < nn) code[first] = 2;
    for (int k = 0; k < nd; ++k) code[dirichlet[k]] = 1;
    m.nn = nn;
    m.E = E;
    m.nnz = (int)col.size();
    m.structured = false;
    CK(up(m.coords, coords, 2 * (size_t)nn, ctx->stream));
    if (E > 0) CK(up(q.elems, elems, 8 * (size_t)E, ctx->stream));
    else CK(q.elems.ensure(32));
    CK(up(m.n2e_ptr, cnt.data(), cnt.size(), ctx->stream));
    if (!n2e.empty()) CK(up(q.n2e, n2e.data(), n2e.size(), ctx->stream));
    else CK(q.n2e.ensure(sizeof(aldic::N2E8)));
    CK(up(m.rowptr, rowptr.data(), rowptr.size(), ctx->stream));
    CK(up(m.col, col.data(), col.size(), ctx->stream));
    CK(up(q.code, code.data(), code.size(), ctx->stream));
    CK(m.aval.ensure(sizeof(double) * col.size()));
    CK(m.mval.ensure(sizeof(double) * col.size()));
    CK(q.cval.ensure(sizeof(double) * col.size()));
    CK(m.diag.ensure(sizeof(double) * nn));
    CK(m.gvec.ensure(sizeof(double) * 2 * nn));
    CK(natgrid_setup(q.nat, coords, nn, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    m.beta = m.mu = m.alpha = -1;
    ctx->have_quad = true;
    return ALDIC_OK;
}

int aldic_subpb2_quadtree(aldic_ctx *ctx, double beta, double mu, const double *U, const double *F,
                          double rtol, int maxit, double *Uhat, double *stats)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_quad) return fail(ctx, ALDIC_ERR_STATE, "subpb2_quadtree: no quadtree mesh");
    if (!U || !F || !Uhat || maxit < 1 || !(rtol > 0))
        return fail(ctx, ALDIC_ERR_ARG, "subpb2_quadtree: bad arguments");
    CK(hipSetDevice(ctx->device));
    aldic::QuadDev &q = ctx->quad;
    const size_t nn = q.m.nn;
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], U, 2 * nn, ctx->stream));
    CK(up(sc[1], F, 4 * nn, ctx->stream));
    CK(sc[4].ensure(sizeof(double) * 2 * nn));
    CK(sc[5].ensure(sizeof(double) * 2 * nn));
    if (q.m.beta != beta || q.m.mu != mu) CK(aldic::quad_assemble(q, beta, mu, ctx->stream));
    CK(aldic::quad_rhs(q, beta, mu, sc[0].as<double>(), sc[1].as<double>(), sc[4].as<double>(),
                       ctx->stream));
    static thread_local DBuf work[8];
    double st[2] = {0, 0};
    CK(aldic::fem_pcg(q.m, sc[4].as<double>(), sc[5].as<double>(), rtol, maxit, work, st,
                      ctx->stream, 8));
    CK(aldic::quad_finish(q, sc[0].as<double>(), sc[5].as<double>(), ctx->stream));
    CK(down(Uhat, sc[5].p, 2 * nn, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (stats) { stats[0] = st[0]; stats[1] = st[1]; }
    return ALDIC_OK;
}

int aldic_nodal_strain_quadtree(aldic_ctx *ctx, const double *U, const int *tri, const double *w,
                                double *F, double *Fgp, double *gpxy)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_quad) return fail(ctx, ALDIC_ERR_STATE, "nodal_strain_quadtree: no quadtree mesh");
    if (!U || (!tri) != (!w) || (tri && !F) || (!F && !Fgp && !gpxy))
        return fail(ctx, ALDIC_ERR_ARG, "nodal_strain_quadtree: bad arguments");
    aldic::QuadDev &q = ctx->quad;
    const size_t nn = q.m.nn, ngp = 4 * (size_t)q.m.E;
    if (tri)
        for (size_t k = 0; k < 3 * nn; ++k)
            if (tri[k] < 0 || (size_t)tri[k] >= ngp)
                return fail(ctx, ALDIC_ERR_ARG, "nodal_strain_quadtree: Gauss point index");
    CK(hipSetDevice(ctx->device));
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], U, 2 * nn, ctx->stream));
    CK(sc[1].ensure(sizeof(double) * 4 * (ngp ? ngp : 1)));
    CK(sc[2].ensure(sizeof(double) * 2 * (ngp ? ngp : 1)));
    CK(aldic::quad_gp_strain(q, sc[0].as<double>(), sc[1].as<double>(), sc[2].as<double>(),
                             ctx->stream));
    if (tri) {
        CK(up(sc[3], tri, 3 * nn, ctx->stream));
        CK(up(sc[4], w, 3 * nn, ctx->stream));
        CK(sc[5].ensure(sizeof(double) * 4 * nn));
        CK(aldic::quad_gp_gather((int)nn, sc[3].as<int>(), sc[4].as<double>(), sc[1].as<double>(),
                                 sc[5].as<double>(), ctx->stream));
        CK(down(F, sc[5].p, 4 * nn, ctx->stream));
    }
    if (Fgp && ngp) CK(down(Fgp, sc[1].p, 4 * ngp, ctx->stream));
    if (gpxy && ngp) CK(down(gpxy, sc[2].p, 2 * ngp, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_nodal_strain(aldic_ctx *ctx, const double *U, int swap_fix, double *F)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_mesh) return fail(ctx, ALDIC_ERR_STATE, "nodal_strain: no mesh");
    if (!U || !F) return fail(ctx, ALDIC_ERR_ARG, "nodal_strain: NULL");
    CK(hipSetDevice(ctx->device));
    const size_t nn = ctx->mesh.nn;
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], U, 2 * nn, ctx->stream));
    CK(sc[1].ensure(sizeof(double) * 4 * nn));
    CK(fem_strain(ctx->mesh, sc[0].as<double>(), swap_fix, sc[1].as<double>(), ctx->stream));
    CK(down(F, sc[1].p, 4 * nn, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

// Runs an outlier/in-painting step on the context's work buffers and maps its status.
static int outlier_status(aldic_ctx *ctx, int r, const char *msg, const char *what)
{
    if (r == 0) return ALDIC_OK;
    if (r == -1) return fail(ctx, ALDIC_ERR_HIP, std::string(what) + ": HIP error");
    return fail(ctx, r, msg ? msg : what);
}

int aldic_inpaint_nans(aldic_ctx *ctx, double *A, int rows, int cols, int method)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!A || rows < 1 || cols < 1 || (method != 3 && method != 4))
        return fail(ctx, ALDIC_ERR_ARG, "inpaint_nans: bad arguments (method 3 or 4)");
    CK(hipSetDevice(ctx->device));
    if (!ctx->outl) ctx->outl = outlier_work_new();
    const size_t N = (size_t)rows * cols;
    DBuf &d = ctx->scratch[0];
    CK(up(d, A, N, ctx->stream));
    const char *msg = nullptr;
    const int r = inpaint_view(*ctx->outl, d.as<double>(), rows, cols, 1, rows, method,
                               ctx->stream, msg);
    if (r) return outlier_status(ctx, r, msg, "inpaint_nans");
    CK(down(A, d.p, N, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_remove_outliers(aldic_ctx *ctx, double *u, double *v, int rows, int cols, double thr,
                          int *flagged, int *n_flagged)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!u || !v || rows < 1 || cols < 1 || !(thr > 0))
        return fail(ctx, ALDIC_ERR_ARG, "remove_outliers: bad arguments");
    CK(hipSetDevice(ctx->device));
    if (!ctx->outl) ctx->outl = outlier_work_new();
    const size_t N = (size_t)rows * cols;
    DBuf *sc = ctx->scratch;
    CK(sc[0].ensure(sizeof(double) * 2 * N));
    CK(hipMemcpyAsync(sc[0].p, u, sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
    CK(hipMemcpyAsync(sc[0].as<double>() + N, v, sizeof(double) * N, hipMemcpyHostToDevice,
                      ctx->stream));
    if (flagged) CK(sc[1].ensure(sizeof(int) * N));
    const char *msg = nullptr;
    int nf = 0;
    const int r = remove_outliers_views(*ctx->outl, sc[0].as<double>(), sc[0].as<double>() + N,
                                        rows, cols, 1, rows, (float)thr, true,
                                        flagged ? sc[1].as<int>() : nullptr, 1, rows, &nf,
                                        ctx->stream, msg);
    if (r) return outlier_status(ctx, r, msg, "remove_outliers");
    CK(hipMemcpyAsync(u, sc[0].p, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipMemcpyAsync(v, sc[0].as<double>() + N, sizeof(double) * N, hipMemcpyDeviceToHost,
                      ctx->stream));
    if (flagged) CK(down(flagged, sc[1].p, N, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (n_flagged) *n_flagged = nf;
    return ALDIC_OK;
}

int aldic_compute_strain(aldic_ctx *ctx, int M, int N, double h, const double *U,
                         const double *F, int method, int rad, int strain_type, double *Ffinite,
                         double *Fworld)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (M < 1 || N < 1 || !(h > 0) || method < 0 || method > 3 || strain_type < 0 ||
        strain_type > 2 || !Ffinite)
        return fail(ctx, ALDIC_ERR_ARG, "compute_strain: bad arguments");
    if ((method == 0 && !F) || (method != 0 && !U))
        return fail(ctx, ALDIC_ERR_ARG, "compute_strain: NULL input");
    // the trimming ring: 0 for the solved F, 1 for finite differences and the FE strains,
    // the plane-fit half width for plane fits (ComputeStrain.m:17,32,45,82)
    const int r = method == 0 ? 0 : (method == 2 ? rad : 1);
    if (method == 2 && (rad < 1 || 2 * rad + 1 > M || 2 * rad + 1 > N))
        return fail(ctx, ALDIC_ERR_ARG, "compute_strain: plane-fit half width must be 1..(min(M,N)-1)/2");
    const size_t nn = (size_t)M * N;
    if (method == 3 && (!ctx->have_mesh || ctx->mesh.nn != (int)nn))
        return fail(ctx, ALDIC_ERR_STATE, "compute_strain: method 3 needs the M x N mesh (aldic_set_mesh)");
    CK(hipSetDevice(ctx->device));
    DBuf *sc = ctx->scratch;
    CK(sc[1].ensure(sizeof(double) * 4 * nn));
    double *Fd = sc[1].as<double>();
    if (method == 0) {
        CK(up(sc[1], F, 4 * nn, ctx->stream));
    } else {
        CK(up(sc[0], U, 2 * nn, ctx->stream));
        if (method == 3)
            CK(fem_strain(ctx->mesh, sc[0].as<double>(), 0, Fd, ctx->stream));
        else
            CK(launch_strain_grad(M, N, h, method, rad, sc[0].as<double>(), Fd, ctx->stream));
    }
    const long nw = (M - 2L * r) > 0 && (N - 2L * r) > 0 ? (M - 2L * r) * (N - 2L * r) : 0;
    CK(sc[2].ensure(sizeof(double) * 4 * (nw ? nw : 1)));
    CK(launch_strain_finish(M, N, strain_type, r, Fd, sc[2].as<double>(), ctx->stream));
    CK(down(Ffinite, Fd, 4 * nn, ctx->stream));
    if (Fworld && nw) CK(down(Fworld, sc[2].p, 4 * (size_t)nw, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_gridfit(aldic_ctx *ctx, int gx, int gy, const double *xnodes, const double *ynodes,
                  const double *z, double rtol, int maxit, double *out, double *stats)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (gx < 2 || gy < 2 || !xnodes || !ynodes || !z || !out || !(rtol > 0) || maxit < 1)
        return fail(ctx, ALDIC_ERR_ARG, "gridfit: bad arguments");
    for (int i = 0; i + 1 < gx; ++i)
        if (!(xnodes[i + 1] > xnodes[i])) return fail(ctx, ALDIC_ERR_ARG, "gridfit: xnodes not increasing");
    for (int j = 0; j + 1 < gy; ++j)
        if (!(ynodes[j + 1] > ynodes[j])) return fail(ctx, ALDIC_ERR_ARG, "gridfit: ynodes not increasing");
    const size_t n = (size_t)gx * gy;
    std::vector<double> u(2 * n, 0.0);  // (z, 0) pairs: the solver fits two components
    std::vector<int> bad(n);
    double sum = 0.0;
    size_t cnt = 0;
    for (size_t i = 0; i < n; ++i)
        if (std::isfinite(z[i])) { sum += z[i]; ++cnt; }
    if (cnt < 3) return fail(ctx, ALDIC_ERR_ARG, "gridfit: insufficient data (fewer than 3 points)");
    for (size_t i = 0; i < n; ++i) {
        bad[i] = std::isfinite(z[i]) ? 0 : 1;
        u[2 * i] = bad[i] ? sum / cnt : z[i];  // starting guess at the empty nodes
    }
    CK(hipSetDevice(ctx->device));
    CK(gridfit_setup(ctx->gf, xnodes, gx, ynodes, gy, ctx->stream));
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], u.data(), 2 * n, ctx->stream));
    CK(up(sc[1], bad.data(), n, ctx->stream));
    CK(sc[2].ensure(4 * sizeof(double)));
    static thread_local DBuf work[8];
    CK(gridfit_solve(ctx->gf, sc[1].as<int>(), sc[0].as<double>(), 1, rtol, maxit, work,
                     sc[2].as<double>(), ctx->stream, 8));
    CK(down(u.data(), sc[0].p, 2 * n, ctx->stream));
    CK(down(stats, sc[2].as<double>() + 1, 2, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < n; ++i) out[i] = u[2 * i];
    return ALDIC_OK;
}

int aldic_natural_fill(aldic_ctx *ctx, int nn, const double *coords, const int *bad, double *U,
                       double *F)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (nn < 0 || (nn > 0 && (!coords || !bad || !U)))
        return fail(ctx, ALDIC_ERR_ARG, "natural_fill: bad arguments");
    if (nn == 0) return ALDIC_OK;
    CK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DBuf *sc = ctx->scratch;
    CK(natgrid_setup(ctx->nat, coords, nn, st));
    CK(up(sc[0], coords, 2 * (size_t)nn, st));
    CK(up(sc[1], bad, (size_t)nn, st));
    CK(up(sc[2], U, 2 * (size_t)nn, st));
    if (F) CK(up(sc[3], F, 4 * (size_t)nn, st));
    CK(natural_fill(ctx->nat, sc[0].as<double>(), sc[1].as<int>(), sc[2].as<double>(),
                    F ? sc[3].as<double>() : nullptr, st));
    CK(down(U, sc[2].p, 2 * (size_t)nn, st));
    if (F) CK(down(F, sc[3].p, 4 * (size_t)nn, st));
    CK(hipStreamSynchronize(st));
    const int nerr = natural_errors(ctx->nat, st);
    if (nerr != 0) {
        char buf[128];
        snprintf(buf, sizeof buf, "natural_fill: %d Sibson cells could not be formed", nerr);
        return fail(ctx, ALDIC_ERR_NUMERIC, buf);
    }
    return ALDIC_OK;
}

int aldic_bicubic(aldic_ctx *ctx, const double *X, const double *Y, long n, double *Z)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "bicubic: no images");
    if (n < 0 || (n > 0 && (!X || !Y || !Z))) return fail(ctx, ALDIC_ERR_ARG, "bicubic: bad arguments");
    if (n == 0) return ALDIC_OK;
    CK(hipSetDevice(ctx->device));
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], X, n, ctx->stream));
    CK(up(sc[1], Y, n, ctx->stream));
    CK(sc[2].ensure(sizeof(double) * n));
    CK(launch_bicubic(ctx->g.as<float>(), ctx->nx, ctx->ny, sc[0].as<double>(), sc[1].as<double>(),
                      n, sc[2].as<double>(), ctx->stream));
    CK(down(Z, sc[2].p, n, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

// ---- learned initial guess -------------------------------------------------------------
long aldic_cnn_nparams(void) { return cnn_nparams(); }

static int cnn_ready(aldic_ctx *ctx, const char *what, bool need_w)
{
    if (!ctx->cnn) ctx->cnn = cnn_work_new();
    if (!ctx->cnn) return fail(ctx, ALDIC_ERR_HIP, what);
    if (need_w && !cnn_have_weights(*ctx->cnn)) {
        std::string m = std::string(what) + ": no network weights";
        return fail(ctx, ALDIC_ERR_STATE, m.c_str());
    }
    return ALDIC_OK;
}

int aldic_cnn_set_weights(aldic_ctx *ctx, const float *w, long n)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!w || n != cnn_nparams()) return fail(ctx, ALDIC_ERR_ARG, "cnn_set_weights: bad arguments");
    CK(hipSetDevice(ctx->device));
    if (int r = cnn_ready(ctx, "cnn_set_weights", false)) return r;
    CK(cnn_set_weights(*ctx->cnn, w, ctx->stream));
    return ALDIC_OK;
}

int aldic_cnn_get_weights(aldic_ctx *ctx, float *w, long n)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!w || n != cnn_nparams()) return fail(ctx, ALDIC_ERR_ARG, "cnn_get_weights: bad arguments");
    CK(hipSetDevice(ctx->device));
    if (int r = cnn_ready(ctx, "cnn_get_weights", true)) return r;
    CK(cnn_get_weights(*ctx->cnn, w, ctx->stream));
    return ALDIC_OK;
}

int aldic_cnn_predict(aldic_ctx *ctx, int S, const double *coords, double *P)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "cnn_predict: no images");
    if (S < 0 || (S > 0 && (!coords || !P))) return fail(ctx, ALDIC_ERR_ARG, "cnn_predict: bad arguments");
    if (S == 0) return ALDIC_OK;
    CK(hipSetDevice(ctx->device));
    if (int r = cnn_ready(ctx, "cnn_predict", true)) return r;
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], coords, 2 * (size_t)S, ctx->stream));
    CK(sc[1].ensure(sizeof(double) * 6 * (size_t)S));
    CK(cnn_forward(*ctx->cnn, ctx->f.as<float>(), ctx->g.as<float>(), ctx->nx, ctx->ny,
                   sc[0].as<double>(), cnn_plan(coords, S), S, sc[1].as<double>(), ctx->stream));
    CK(down(P, sc[1].p, 6 * (size_t)S, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_cnn_add_samples(aldic_ctx *ctx, int S, const double *coords, const double *P)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "cnn_add_samples: no images");
    if (S < 0 || (S > 0 && (!coords || !P))) return fail(ctx, ALDIC_ERR_ARG, "cnn_add_samples: bad arguments");
    if (S == 0) return ALDIC_OK;
    CK(hipSetDevice(ctx->device));
    if (int r = cnn_ready(ctx, "cnn_add_samples", false)) return r;
    DBuf *sc = ctx->scratch;
    CK(up(sc[0], coords, 2 * (size_t)S, ctx->stream));
    CK(up(sc[1], P, 6 * (size_t)S, ctx->stream));
    CK(cnn_add_samples(*ctx->cnn, ctx->f.as<float>(), ctx->g.as<float>(), ctx->nx, ctx->ny,
                       sc[0].as<double>(), sc[1].as<double>(), S, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_cnn_clear_samples(aldic_ctx *ctx)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (ctx->cnn) cnn_clear_samples(*ctx->cnn);
    return ALDIC_OK;
}

int aldic_cnn_train(aldic_ctx *ctx, int steps, int batch, float lr, unsigned long long seed,
                    double *loss)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (steps < 0 || batch <= 0 || batch > 65536 || !(lr > 0.f) || (steps > 0 && !loss))
        return fail(ctx, ALDIC_ERR_ARG, "cnn_train: bad arguments");
    CK(hipSetDevice(ctx->device));
    if (int r = cnn_ready(ctx, "cnn_train", true)) return r;
    const long n = cnn_n_samples(*ctx->cnn);
    if (n <= 0) return fail(ctx, ALDIC_ERR_STATE, "cnn_train: empty training set");
    if (n >= (1L << 31)) return fail(ctx, ALDIC_ERR_ARG, "cnn_train: training set too large");
    DBuf *sc = ctx->scratch;
    CK(sc[2].ensure(sizeof(double) * (size_t)std::max(steps, 1)));
    try {
        // batches drawn with replacement by splitmix64 from the seed
        std::vector<int> idx((size_t)batch);
        unsigned long long st = seed;
        for (int it = 0; it < steps; ++it) {
            for (int b = 0; b < batch; ++b) {
                st += 0x9e3779b97f4a7c15ull;
                unsigned long long z = st;
                z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
                z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
                z ^= z >> 31;
                idx[(size_t)b] = (int)(z % (unsigned long long)n);
            }
            CK(cnn_step(*ctx->cnn, idx.data(), batch, lr, sc[2].as<double>() + it, ctx->stream));
            // the index vector is reused: the copy above must finish before it changes
            CK(hipStreamSynchronize(ctx->stream));
        }
    } catch (...) {
        return fail(ctx, ALDIC_ERR_HIP, "cnn_train: host allocation failed");
    }
    if (steps > 0) CK(down(loss, sc[2].p, (size_t)steps, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_cnn_grad(aldic_ctx *ctx, int B, const int *rows, float *grad, double *loss,
                   double *outputs)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (B <= 0 || !rows || !grad || !loss) return fail(ctx, ALDIC_ERR_ARG, "cnn_grad: bad arguments");
    CK(hipSetDevice(ctx->device));
    if (int r = cnn_ready(ctx, "cnn_grad", true)) return r;
    const long n = cnn_n_samples(*ctx->cnn);
    for (int b = 0; b < B; ++b)
        if (rows[b] < 0 || rows[b] >= n) return fail(ctx, ALDIC_ERR_ARG, "cnn_grad: row outside the training set");
    DBuf *sc = ctx->scratch;
    CK(sc[2].ensure(sizeof(double)));
    CK(cnn_step(*ctx->cnn, rows, B, 0.f, sc[2].as<double>(), ctx->stream));
    CK(hipMemcpyAsync(grad, cnn_grad(*ctx->cnn), sizeof(float) * cnn_nparams(), hipMemcpyDeviceToHost,
                      ctx->stream));
    CK(down(loss, sc[2].p, 1, ctx->stream));
    if (outputs) {
        std::vector<float> y((size_t)B * 6);
        CK(hipMemcpyAsync(y.data(), cnn_outputs(*ctx->cnn), sizeof(float) * y.size(), hipMemcpyDeviceToHost,
                          ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i < y.size(); ++i) outputs[i] = y[i];
    }
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

// ---- LocalICGN from the search or from the learned start -------------------------------
// The start and 6-DOF IC-GN of LocalICGN over the whole-field search grid, resident on
// the device: start 0 = the integer search (IntegerSearch.m:54 -> LocalICGN.m:91-117),
// start 1 = the learned warp (aldic_cnn_predict on the grid's nodes) with the FP32
// search and IC-GN from its peak redone on the subsets whose IC-GN from the warp did not
// converge within max_it residual evaluations or ended below min_zncc.
int aldic_local_icgn(aldic_ctx *ctx, int start, const int *grid, int ws, int step, const int *s, double tol,
                     int max_it, double min_zncc, double *x, double *y, double *U, double *F, int *iters,
                     double *zncc, int *fallback, int *n_fallback, float *times_ms)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "local_icgn: no images");
    if ((start != 0 && start != 1) || !grid || !s || !U || !F || !iters || !(tol > 0) || max_it < 1 ||
        s[0] < 0 || s[1] < 0)
        return fail(ctx, ALDIC_ERR_ARG, "local_icgn: bad arguments");
    int gr[4];
    if (aldic_search_grid(ctx->nx, ctx->ny, grid, ws, step, s, gr))
        return fail(ctx, ALDIC_ERR_ARG, "local_icgn: bad grid arguments");
    const int S = gr[2] * gr[3];
    if (n_fallback) *n_fallback = 0;
    if (S <= 0) return ALDIC_OK;
    const int ppl = icgn_ppl_for(ws);
    if (ppl == 0 || ws < 2) return fail(ctx, ALDIC_ERR_ARG, "local_icgn: subset size must be 2..64");
    if (!search_fits(ctx, start == 0 && ctx->images_from_raw, ws, s[0], s[1]))
        return fail(ctx, ALDIC_ERR_ARG, "local_icgn: window too large for LDS");
    CK(hipSetDevice(ctx->device));
    if (start == 1)
        if (int r = cnn_ready(ctx, "local_icgn", true)) return r;
    hipStream_t st = ctx->stream;
    DBuf *w = ctx->lic;
    // 0: coords (2S), 1: U0 (2S) then the search's x, y, u, v, cc (5S), 2: P (6S), 3: F0 (4S),
    // 4: U, F (6S), 5: H21 (21S), 6: zncc (S) + iters (S ints), 7: flags, list, count, scratch
    CK(w[0].ensure(sizeof(double) * 2 * (size_t)S));
    CK(w[1].ensure(sizeof(double) * 7 * (size_t)S + sizeof(int) * (size_t)S));
    CK(w[2].ensure(sizeof(double) * 6 * (size_t)S));
    CK(w[3].ensure(sizeof(double) * 4 * (size_t)S));
    CK(w[4].ensure(sizeof(double) * 6 * (size_t)S));
    CK(w[5].ensure(sizeof(double) * 21 * (size_t)S));
    CK(w[6].ensure(sizeof(double) * (size_t)S + sizeof(int) * (size_t)S));
    // flags (S), fallback list (S), its count (2), list-building runs, -2 list (S + 1)
    const size_t nrun = ((size_t)S + 4095) / 4096;
    CK(w[7].ensure(sizeof(int) * (3 * (size_t)S + 3 + nrun)));
    double *xy = w[0].as<double>(), *U0 = w[1].as<double>(), *so = U0 + 2 * (size_t)S;
    double *Ud = w[4].as<double>(), *Fd = Ud + 2 * (size_t)S, *zd = w[6].as<double>();
    int *itd = reinterpret_cast<int *>(zd + S), *flag = w[7].as<int>(), *list = flag + S, *cnt = list + S;
    int *runs = cnt + 2, *rl = runs + nrun;
    // the node centres on the host and the network's plan, before the timed span
    std::vector<double> xyh(2 * (size_t)S);
    for (int i = 0; i < S; ++i) {
        xyh[2 * i] = (2.0 * (gr[0] + (i % gr[2]) * step) + ws) / 2.0;
        xyh[2 * i + 1] = (2.0 * (gr[1] + (i / gr[2]) * step) + ws) / 2.0;
    }
    const CnnPlan plan = start == 1 ? cnn_plan(xyh.data(), S) : CnnPlan{};
    hipEvent_t *ev = ctx->ev;
    CK(hipEventRecord(ev[0], st));
    CK(launch_grid_coords(gr[0], gr[1], gr[2], gr[3], step, ws, xy, st));
    SearchArgs sa{};
    sa.f = ctx->f.as<float>();
    sa.g = ctx->g.as<float>();
    sa.nx = ctx->nx;
    sa.ny = ctx->ny;
    sa.x1 = gr[0];
    sa.y1 = gr[1];
    sa.Nx = gr[2];
    sa.Ny = gr[3];
    sa.ws = ws;
    sa.step = step;
    sa.sx = s[0];
    sa.sy = s[1];
    sa.x = so;
    sa.y = so + S;
    sa.u = so + 2 * (size_t)S;
    sa.v = so + 3 * (size_t)S;
    sa.ccmax = so + 4 * (size_t)S;
    sa.status = reinterpret_cast<int *>(so + 5 * (size_t)S);
    sa.U0 = U0;
    const double *F0 = nullptr;
    if (start == 0) {
        search_raw8(ctx, sa);
        CK(launch_search(sa, st));
    } else {
        CK(cnn_forward(*ctx->cnn, ctx->f.as<float>(), ctx->g.as<float>(), ctx->nx, ctx->ny, xy, plan, S,
                       w[2].as<double>(), st));
        CK(launch_split_start(w[2].as<double>(), S, U0, w[3].as<double>(), st));
        F0 = w[3].as<double>();
    }
    CK(hipEventRecord(ev[1], st));
    IcgnArgs a{};
    a.f = ctx->f.as<float>();
    a.g = ctx->g.as<float>();
    a.gx = ctx->gx.as<float>();
    a.gy = ctx->gy.as<float>();
    a.nx = ctx->nx;
    a.ny = ctx->ny;
    a.S = S;
    a.coords = xy;
    a.ws_all = ws;
    a.tol = tol;
    a.U0 = U0;
    a.F0 = F0;
    a.U = Ud;
    a.F = Fd;
    a.H21 = w[5].as<double>();
    a.zncc = zd;
    a.iters = itd;
    a.nlist = S;
    CK(launch_icgn6(a, ppl, st));
    // windows the mask rule enlarged beyond the class (status -2) run one class up
    const int ppl_big = icgn_ppl_for((int)std::floor(std::sqrt(2.0) * ws + 0.5));
    if (ppl_big != 0 && ppl_big != ppl) {
        IcgnArgs ar = a;
        CK(launch_list_eq(itd, S, -2, rl, rl + S, runs, st));
        ar.list = rl;
        ar.nlist_dev = rl + S;
        CK(launch_icgn6(ar, ppl_big, st));
    }
    CK(hipEventRecord(ev[2], st));
    int nfb = 0;
    if (start == 1) {
        // fallback: FP32 search on the flagged subsets, IC-GN from its peaks there
        // (the list's length comes back to the host so that the search launches one
        // workgroup per listed subset only)
        CK(launch_flag_fallback(itd, zd, S, max_it, min_zncc, flag, st));
        CK(launch_list_eq(flag, S, 1, list, cnt, runs, st));
        CK(hipMemcpyAsync(&nfb, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        SearchArgs sl = sa;
        sl.list = list;
        sl.nlist_dev = cnt;
        CK(launch_search_list(sl, nfb, st));
        CK(hipEventRecord(ev[3], st));
        if (nfb > 0) {
            IcgnArgs af = a;
            af.F0 = nullptr;
            af.list = list;
            af.nlist = nfb;
            CK(launch_icgn6(af, ppl, st));
            if (ppl_big != 0 && ppl_big != ppl) {
                IcgnArgs ar = a;
                ar.F0 = nullptr;
                CK(launch_list_eq(itd, S, -2, rl, rl + S, runs, st));
                ar.list = rl;
                ar.nlist_dev = rl + S;
                CK(launch_icgn6(ar, ppl_big, st));
            }
        }
        CK(hipEventRecord(ev[4], st));
    } else {
        CK(hipEventRecord(ev[3], st));
        CK(hipEventRecord(ev[4], st));
    }
    CK(hipStreamSynchronize(st));
    if (n_fallback) *n_fallback = nfb;
    if (times_ms) {
        for (int k = 0; k < 4; ++k) CK(hipEventElapsedTime(&times_ms[k], ev[k], ev[k + 1]));
        CK(hipEventElapsedTime(&times_ms[4], ev[0], ev[4]));
    }
    CK(down(U, Ud, 2 * (size_t)S, st));
    CK(down(F, Fd, 4 * (size_t)S, st));
    CK(down(iters, itd, S, st));
    if (zncc) CK(down(zncc, zd, S, st));
    if (fallback) {
        if (start == 1) CK(down(fallback, flag, S, st));
    }
    CK(hipStreamSynchronize(st));
    if (fallback && start == 0) std::fill(fallback, fallback + S, 0);
    for (int i = 0; i < S; ++i) {
        if (x) x[i] = xyh[2 * i];
        if (y) y[i] = xyh[2 * i + 1];
    }
    return ALDIC_OK;
}

// ---- device-resident frame ------------------------------------------------------------
int aldic_frame_stage(aldic_ctx *ctx, const void *f, const void *g, int nx, int ny, int dtype,
                      const uint8_t *mask, const int *roi)
{
    if (!ctx) return ALDIC_ERR_ARG;
    CK(hipSetDevice(ctx->device));
    int r = stage_raw(ctx, f, g, nx, ny, dtype, mask, roi);
    if (r) return r;
    CK(hipStreamSynchronize(ctx->stream));
    ctx->fr.ready = false;
    return ALDIC_OK;
}

int aldic_frame_setup(aldic_ctx *ctx, const aldic_frame_opts *o, int *n_nodes)
{
    if (!ctx || !o) return ALDIC_ERR_ARG;
    if (!ctx->rawf.p) return fail(ctx, ALDIC_ERR_STATE, "frame_setup: no staged images");
    if (o->n_admm < 0 || o->cg_maxit < 1 || !(o->cg_rtol > 0))
        return fail(ctx, ALDIC_ERR_ARG, "frame_setup: bad options");
    if (icgn_ppl_for(o->ws) == 0) return fail(ctx, ALDIC_ERR_ARG, "frame_setup: ws must be 2..64");
    if (!search_fits(ctx, true, o->ws, o->s[0], o->s[1]))
        return fail(ctx, ALDIC_ERR_ARG, "frame_setup: search window too large for LDS");
    CK(hipSetDevice(ctx->device));
    FrameState &F = ctx->fr;
    F.o = *o;
    if (!F.ev_skern) CK(hipEventCreate(&F.ev_skern));
    if (!F.ev_ikern) CK(hipEventCreate(&F.ev_ikern));
    for (size_t nev = 8 + 5 * (size_t)(o->n_admm + 1); F.ev.size() < nev;) {
        hipEvent_t e;
        CK(hipEventCreate(&e));
        F.ev.push_back(e);
    }
    int gr[4];
    if (aldic_search_grid(ctx->nx, ctx->ny, o->grid, o->ws, o->step, o->s, gr))
        return fail(ctx, ALDIC_ERR_ARG, "frame_setup: bad grid");
    F.x1 = gr[0];
    F.y1 = gr[1];
    F.Nx = gr[2];
    F.Ny = gr[3];
    F.S = F.Nx * F.Ny;
    if (F.Nx < 2 || F.Ny < 2) return fail(ctx, ALDIC_ERR_ARG, "frame_setup: fewer than 2x2 subsets");
    const int S = F.S;
    // MeshSetUp.m:23-39 -- node (k, l) at the subset centre, x fastest; Q4 elements
    F.hcoords.resize(2 * (size_t)S);
    for (int l = 0; l < F.Ny; ++l)
        for (int k = 0; k < F.Nx; ++k) {
            F.hcoords[2 * ((size_t)l * F.Nx + k)] = F.x1 + k * o->step + 0.5 * o->ws;
            F.hcoords[2 * ((size_t)l * F.Nx + k) + 1] = F.y1 + l * o->step + 0.5 * o->ws;
        }
    F.helems.clear();
    for (int l = 0; l + 1 < F.Ny; ++l)
        for (int k = 0; k + 1 < F.Nx; ++k) {
            const int n0 = l * F.Nx + k;
            F.helems.insert(F.helems.end(), {n0, n0 + 1, n0 + F.Nx + 1, n0 + F.Nx});
        }
    int r = aldic_set_mesh(ctx, F.hcoords.data(), S, F.helems.data(), (int)F.helems.size() / 4);
    if (r) return r;
    CK(up(F.coords, F.hcoords.data(), F.hcoords.size(), ctx->stream));
    CK(natgrid_setup(F.nat, F.hcoords.data(), S, ctx->stream));
    CK(F.nat_err.ensure(sizeof(int)));
    CK(F.nbad.ensure(sizeof(int) * kMaxSteps));
    if (!F.nat_err_h) CK(hipHostMalloc((void **)&F.nat_err_h, sizeof(int), hipHostMallocDefault));
    const size_t sd = sizeof(double) * S;
    DBuf *v2[] = {&F.U0, &F.U6, &F.Ucur, &F.U1, &F.U2, &F.vdual};
    for (DBuf *b : v2) CK(b->ensure(2 * sd));
    DBuf *v4[] = {&F.F6, &F.Fcur, &F.F1, &F.F2, &F.udual};
    for (DBuf *b : v4) CK(b->ensure(4 * sd));
    CK(F.H21.ensure(21 * sd));
    CK(F.zncc.ensure(sd));
    CK(F.it6.ensure(sizeof(int) * S));
    CK(F.it2.ensure(sizeof(int) * S));
    CK(F.bad.ensure(sizeof(int) * S));
    CK(F.part.ensure(3 * 256 * sizeof(double)));
    CK(F.cc.ensure(8 * sizeof(double) * S));
    CK(F.rl.ensure(sizeof(int) * ((size_t)S + 1 + (S + 4095) / 4096)));  // list, count, run counts
    // gridfits per frame: Subpb1's refit per ADMM step, and with smoothing two per field
    // smoothed (F1; F2 of every step; U1, U2 of every ADMM step)
    F.gf_cap = 4 + 8 * (o->n_admm + 1);
    CK(F.gfstat.ensure(3 * sizeof(double) * F.gf_cap));
    F.gf_budget = 16;
    F.gf_reruns = 0;
    F.gf_most = 0;
    F.gf_count = 0;
    if (F.gfh) CK(hipHostFree(F.gfh));
    F.gfh = nullptr;
    CK(hipHostMalloc((void **)&F.gfh, 3 * sizeof(double) * F.gf_cap, hipHostMallocDefault));
    if (!F.hostd) CK(hipHostMalloc((void **)&F.hostd, 16 * sizeof(double), hipHostMallocDefault));
    CK(F.errs.ensure(16 * sizeof(double)));
    if (o->smooth) {
        CK(F.Us.ensure(2 * sd));
        CK(F.Fs.ensure(4 * sd));
    }
    if (o->beta_sweep) CK(F.Fs.ensure(4 * sd));
    if (o->admm_tol > 0) {
        CK(F.U1s.ensure(2 * sd));
        CK(F.U1old.ensure(2 * sd));
        CK(F.U2old.ensure(2 * sd));
    }
    F.beta_used = o->beta;
    F.admm_steps = 0;
    F.upd = F.upd2 = NAN;
    F.n_fallback = -1;
    if (o->learned_start) {
        // the network's warps, F0, and the fallback list (flags, list, count, run counts)
        if (int r = cnn_ready(ctx, "frame_setup: learned_start", true)) return r;
        if (!search_fits(ctx, false, o->ws, o->s[0], o->s[1]))
            return fail(ctx, ALDIC_ERR_ARG, "frame_setup: fallback search window too large for LDS");
        F.plan = cnn_plan(F.hcoords.data(), S);
        CK(F.P6.ensure(6 * sd));
        CK(F.F0.ensure(4 * sd));
        CK(F.lflag.ensure(sizeof(int) * (2 * (size_t)S + 2 + ((size_t)S + 4095) / 4096)));
        if (!F.nfb_h) CK(hipHostMalloc((void **)&F.nfb_h, sizeof(int), hipHostMallocDefault));
    }
    CK(F.F2b.ensure(sizeof(double) * 4 * S));
    CK(F.tmp.ensure(2 * sd));
    CK(ctx->scratch[5].ensure(2 * sd + 64));
    CK(fem_assemble(ctx->mesh, o->beta, o->mu, 0.0, ctx->stream));
    if (!o->no_subpb1_fit || o->smooth) {
        // gridfit grid (Subpb1's refit and the smoothing passes): xnodes = unique node x,
        // ynodes = unique node y (= min:step:max of funSmoothDisp.m on this mesh)
        if (F.Nx < 2 || F.Ny < 2)
            return fail(ctx, ALDIC_ERR_ARG, "frame_setup: gridfit needs at least 2 x 2 nodes");
        std::vector<double> xn(F.Nx), yn(F.Ny);
        for (int k = 0; k < F.Nx; ++k) xn[k] = F.hcoords[2 * (size_t)k];
        for (int l = 0; l < F.Ny; ++l) yn[l] = F.hcoords[2 * (size_t)l * F.Nx + 1];
        CK(gridfit_setup(F.gf, xn.data(), F.Nx, yn.data(), F.Ny, ctx->stream));
    }
    CK(hipStreamSynchronize(ctx->stream));
    F.ready = true;
    if (n_nodes) *n_nodes = S;
    return ALDIC_OK;
}

// One pass of the frame.  gf_poll false: each Subpb1 gridfit runs a fixed budget of
// iteration kernels with no host wait and records whether it converged in F.gfh; true:
// each polls the host for convergence.
static int frame_pass(aldic_ctx *ctx, bool gf_poll, bool seeded)
{
    FrameState &F = ctx->fr;
    hipStream_t st = ctx->stream;
    const aldic_frame_opts &o = F.o;
    const int S = F.S;
    static thread_local DBuf work[8];
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // Stage boundaries are events in one pool; the host waits once, at the end, so the
    // stream never drains between stages.  spans: (stage, first event, second event).
    std::vector<hipEvent_t> &ev = F.ev;
    int ne = 0;
    std::vector<int> spans;
    auto mark = [&]() -> int {
        if (ne >= (int)ev.size()) return -1;
        return hipEventRecord(ev[ne], st) == hipSuccess ? ne++ : -1;
    };
    // boundaries other than the search's, the 6-DOF IC-GN's and the frame's only with
    // stage_times
    auto smark = [&]() -> int { return o.stage_times ? mark() : -1; };
    auto span = [&](int stage, int e0, int e1) {
        spans.push_back(stage); spans.push_back(e0); spans.push_back(e1);
    };
    const int e_start = mark();
    CK(natural_err_clear(F.nat_err.as<int>(), st));
    CK(launch_iset(F.nbad.as<int>(), kMaxSteps, -1, st));
    CK(run_prep(ctx));
    ctx->have_images = true;
    ctx->images_from_raw = true;
    const int e_prep = mark();
    span(0, e_start, e_prep);
    // integer search -> U0
    SearchArgs sa{};
    sa.f = ctx->f.as<float>();
    sa.g = ctx->g.as<float>();
    sa.nx = ctx->nx;
    sa.ny = ctx->ny;
    sa.x1 = F.x1;
    sa.y1 = F.y1;
    sa.Nx = F.Nx;
    sa.Ny = F.Ny;
    sa.ws = o.ws;
    sa.step = o.step;
    sa.sx = o.s[0];
    sa.sy = o.s[1];
    double *sbuf = F.tmp.as<double>();
    CK(ctx->scratch[4].ensure(sizeof(double) * 4 * S + sizeof(int) * S));
    double *s4 = ctx->scratch[4].as<double>();
    sa.x = s4;
    sa.y = s4 + S;
    sa.u = s4 + 2 * (size_t)S;
    sa.v = s4 + 3 * (size_t)S;
    sa.ccmax = sbuf;
    sa.qf = nullptr;
    sa.status = reinterpret_cast<int *>(s4 + 4 * (size_t)S);
    sa.U0 = F.U0.as<double>();
    sa.ev_kernel = F.ev_skern;
    const SearchArgs sa32 = sa;   // the FP32 search (learned start's fallback list)
    search_raw8(ctx, sa);
    // learned start: the network's warp of every node in place of the search
    const bool learned = o.learned_start && !seeded;
    if (learned) {
        CK(cnn_forward(*ctx->cnn, ctx->f.as<float>(), ctx->g.as<float>(), ctx->nx, ctx->ny,
                       F.coords.as<double>(), F.plan, S, F.P6.as<double>(), st));
        CK(launch_split_start(F.P6.as<double>(), S, F.U0.as<double>(), F.F0.as<double>(), st));
        CK(hipEventRecord(F.ev_skern, st));
    } else if (!seeded) {
        CK(launch_search(sa, st));
    } else {
        // main_ALDIC.m:136-137: the previous frame's U is the initial guess (no search)
        CK(hipEventRecord(F.ev_skern, st));
        CK(hipMemcpyAsync(F.U0.p, F.seed.p, 16 * (size_t)S, hipMemcpyDeviceToDevice, st));
    }
    if (!seeded && !learned && o.outlier_thr > 0) {
        // IntegerSearch.m:68-76 (mask holes NaN), :114 (funRemoveOutliers at Thr0) and
        // Init.m:13: flagged and NaN nodes in-painted by springs in one solve
        if (ctx->have_mask)
            CK(launch_mask_nodes(F.U0.as<double>(), F.coords.as<double>(), S,
                                 ctx->mask.as<uint8_t>(), ctx->nx, ctx->ny, st));
        if (!ctx->outl) ctx->outl = outlier_work_new();
        const char *msg = nullptr;
        const int r0 = remove_outliers_views(*ctx->outl, F.U0.as<double>(), F.U0.as<double>() + 1,
                                             F.Ny, F.Nx, 2L * F.Nx, 2L, (float)o.outlier_thr,
                                             false, nullptr, 0, 0, &F.n_outliers, st, msg);
        if (r0) return outlier_status(ctx, r0, msg, "frame: search outlier test");
    }
    // (with outlier_thr 0 the search's NaNs go to LocalICGN as they are: their subsets end
    // non-finite, status 101, and the bad-point refill covers them)
    const int e_search = mark();
    span(1, e_prep, e_search);

    // 6-DOF IC-GN (LocalICGN)
    IcgnArgs a{};
    a.f = ctx->f.as<float>();
    a.g = ctx->g.as<float>();
    a.gx = ctx->gx.as<float>();
    a.gy = ctx->gy.as<float>();
    a.nx = ctx->nx;
    a.ny = ctx->ny;
    a.S = S;
    a.coords = F.coords.as<double>();
    a.ws = nullptr;
    a.ws_all = o.ws;
    a.tol = o.tol;
    a.U0 = F.U0.as<double>();
    a.F0 = learned ? F.F0.as<double>() : nullptr;
    a.U = F.U6.as<double>();
    a.F = F.F6.as<double>();
    a.H21 = F.H21.as<double>();
    a.zncc = F.zncc.as<double>();
    a.iters = F.it6.as<int>();
    a.nlist = S;
    a.cc = F.cc.as<double>();
    a.cc_mode = 1;
    a.Ucopy = F.U1.as<double>();  // refilled in place by the bad-point step
    a.Fcopy = F.F1.as<double>();
    const int ppl = icgn_ppl_for(o.ws);
    CK(launch_icgn6(a, ppl, st));
    CK(hipEventRecord(F.ev_ikern, st));
    // windows the mask rule enlarged beyond the class (status -2, funICGN.m:56-60) run
    // again one class up, over a list built on the device
    const int ppl_big = icgn_ppl_for((int)std::floor(std::sqrt(2.0) * o.ws + 0.5));
    const bool relaunch = ppl_big != 0 && ppl_big != ppl;
    IcgnArgs ar = a;
    if (relaunch) {
        int *rl = F.rl.as<int>();
        CK(launch_list_eq(F.it6.as<int>(), S, -2, rl, rl + S, rl + S + 1, st));
        ar.list = rl;
        ar.nlist = S;
        ar.nlist_dev = rl + S;
        CK(launch_icgn6(ar, ppl_big, st));
    }
    if (learned) {
        // fallback: the subsets whose IC-GN missed status 1..20 or ended with ZNCC below
        // 0.9 are searched (FP32, that list only: its length comes to the host so the
        // search launches one workgroup per listed subset) and run again from the peak
        int *flag = F.lflag.as<int>(), *list = flag + S, *cnt = list + S, *runs = cnt + 2;
        CK(launch_flag_fallback(F.it6.as<int>(), F.zncc.as<double>(), S, 20, 0.9, flag, st));
        CK(launch_list_eq(flag, S, 1, list, cnt, runs, st));
        CK(hipMemcpyAsync(F.nfb_h, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        const int nfb = *F.nfb_h;
        F.n_fallback = nfb;
        if (nfb > 0) {
            SearchArgs sl = sa32;
            sl.ev_kernel = nullptr;
            sl.list = list;
            sl.nlist_dev = cnt;
            CK(launch_search_list(sl, nfb, st));
            IcgnArgs af = a;
            af.F0 = nullptr;
            af.list = list;
            af.nlist = nfb;
            CK(launch_icgn6(af, ppl, st));
            if (relaunch) {
                IcgnArgs ar2 = af;
                int *rl = F.rl.as<int>();
                CK(launch_list_eq(F.it6.as<int>(), S, -2, rl, rl + S, rl + S + 1, st));
                ar2.list = rl;
                ar2.nlist = S;
                ar2.nlist_dev = rl + S;
                CK(launch_icgn6(ar2, ppl_big, st));
            }
        }
    } else {
        F.n_fallback = -1;
    }
    const int e_icgn6 = mark();
    span(2, e_search, e_icgn6);
    // LocalICGN post-processing (U1, F1 hold copies of U6, F6): the bad-point rule
    // (LocalICGN.m:119-139) and the natural-neighbour refill of U and F (:141-161)
    CK(launch_badmark(F.it6.as<int>(), S, 1, nullptr, nullptr, F.U1.as<double>(),
                      F.F1.as<double>(), F.bad.as<int>(), F.part.as<double>(), st));
    CK(natural_fill(F.nat, F.coords.as<double>(), F.bad.as<int>(), F.U1.as<double>(),
                    F.F1.as<double>(), st, F.nat_err.as<int>()));
    const int e_post6 = smark();
    span(3, e_icgn6, e_post6);

    // Global steps.  F1 is F6 refilled for step 1 and the previous F2 afterwards
    // (FSubpb1 = FSubpb2, main_ALDIC.m:364): F2 alternates between two buffers instead of
    // being copied.  Step 1 runs with zero duals (nullptr).
    double *b = ctx->scratch[5].as<double>();
    double *F2buf[2] = {F.F2.as<double>(), F.F2b.as<double>()};
    int cur = 0;
    const double *F1p = F.F1.as<double>();
    double *U1 = F.U1.as<double>(), *U2 = F.U2.as<double>();
    const bool smooth = o.smooth != 0, stop = o.admm_tol > 0.0;
    // every gridfit records {converged, iterations, residual} for the budget check
    int ngf = 0;
    auto gfit = [&](double *X, int np, const int *bad) -> hipError_t {
        double *rec = ngf < F.gf_cap ? F.gfstat.as<double>() + 3 * ngf++ : nullptr;
        const bool poll = gf_poll || !rec;
        return gridfit_solve(F.gf, bad, X, np, 1e-10, poll ? 4000 : F.gf_budget, work, rec, st,
                             poll ? 8 : 0);
    };
    // funSmoothDisp.m / funSmoothStrain.m: two passes of (gridfit of every node, 3x3
    // Gaussian); nc values per node, tmp holds the first pass
    auto smooth_field = [&](double *X, int nc, double *tmp) -> int {
        CK(gfit(X, nc / 2, nullptr));
        CK(launch_blur3(X, tmp, F.Nx, F.Ny, nc, 0.5, st));
        CK(gfit(tmp, nc / 2, nullptr));
        CK(launch_blur3(tmp, X, F.Nx, F.Ny, nc, 0.5, st));
        return ALDIC_OK;
    };
    int r;
    if (smooth && (r = smooth_field(F.F1.as<double>(), 4, F.Fs.as<double>()))) return r;

    double beta = o.beta;
    if (o.beta_sweep) {
        // main_ALDIC.m:237,274-301 (FE branch): the step-1 global step at six candidates
        const double fac[6] = {1e-3, std::sqrt(1e-5), 1e-2, std::sqrt(1e-3), 1e-1, std::sqrt(1e-1)};
        double bl[6], e1[6], e2[6];
        double *errs = F.errs.as<double>();
        for (int k = 0; k < 6; ++k) {
            bl[k] = fac[k] * o.step * o.step * o.mu;
            CK(fem_assemble(ctx->mesh, bl[k], o.mu, 0.0, st));
            CK(fem_rhs(ctx->mesh, bl[k], o.mu, U1, F1p, nullptr, nullptr, b, st));
            CK(fem_solve(ctx->mesh, b, U2, o.cg_rtol, o.cg_maxit, work, nullptr, st, 8));
            CK(fem_strain(ctx->mesh, U2, 0, F.Fs.as<double>(), st));
            CK(launch_sqdiff(U1, U2, 2 * (long)S, errs + 2 * k, st));
            CK(launch_sqdiff(F1p, F.Fs.as<double>(), 4 * (long)S, errs + 2 * k + 1, st));
        }
        CK(hipMemcpyAsync(F.hostd, errs, 12 * sizeof(double), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        for (int k = 0; k < 6; ++k) {
            e1[k] = std::sqrt(F.hostd[2 * k]);
            e2[k] = std::sqrt(F.hostd[2 * k + 1]);
        }
        beta = beta_choice(bl, e1, e2);
        CK(fem_assemble(ctx->mesh, beta, o.mu, 0.0, st));
    }
    F.beta_used = beta;

    auto global_step = [&](bool zero_duals, bool smooth_u) -> int {
        const int g0 = smark();
        CK(fem_rhs(ctx->mesh, beta, o.mu, U1, F1p, zero_duals ? nullptr : F.udual.as<double>(),
                   zero_duals ? nullptr : F.vdual.as<double>(), b, st));
        double gst[2] = {0.0, 0.0};
        CK(fem_solve(ctx->mesh, b, U2, o.cg_rtol, o.cg_maxit, work,
                     ctx->mesh.structured ? nullptr : gst, st, 8));
        if (!ctx->mesh.structured && !(std::isfinite(gst[1]) && gst[1] <= o.cg_rtol))
            return fail(ctx, ALDIC_ERR_NUMERIC, "frame_run: global-step PCG did not converge");
        const int g1 = smark();
        span(4, g0, g1);
        // nodal strain and the dual update of the FEM branch (main_ALDIC.m:340,440):
        // udual = F2 - F1, vdual = U2 - U1 (after the smoothing passes when they run)
        if (!smooth) {
            CK(fem_strain(ctx->mesh, U2, 0, F2buf[cur], st, F1p, U1, F.udual.as<double>(),
                          F.vdual.as<double>()));
        } else {
            int rr;
            if (smooth_u && (rr = smooth_field(U2, 2, F.Us.as<double>()))) return rr;
            CK(fem_strain(ctx->mesh, U2, 0, F2buf[cur], st));
            if ((rr = smooth_field(F2buf[cur], 4, F.Fs.as<double>()))) return rr;
            CK(launch_sub(F2buf[cur], F1p, F.udual.as<double>(), 4 * (long)S, st));
            CK(launch_sub(U2, U1, F.vdual.as<double>(), 2 * (long)S, st));
        }
        span(6, g1, smark());
        return ALDIC_OK;
    };
    // the stop rule compares Subpb1's U (before smoothing; LocalICGN's for step 1) and U2
    // with the previous step's
    if (stop) CK(hipMemcpyAsync(F.U1old.p, U1, 16 * (size_t)S, hipMemcpyDeviceToDevice, st));
    r = global_step(true, false);
    if (r) return r;
    if (stop) CK(hipMemcpyAsync(F.U2old.p, U2, 16 * (size_t)S, hipMemcpyDeviceToDevice, st));
    F.admm_steps = 0;
    F.upd = F.upd2 = NAN;

    for (int step = 0; step < o.n_admm; ++step) {
        // Subpb1: 2-DOF IC-GN from (U2, F2)
        const int i0 = smark();
        IcgnArgs a2 = a;
        a2.U0 = nullptr;
        a2.Hin = F.H21.as<double>();
        a2.vdual = F.vdual.as<double>();
        a2.Uold = U2;
        a2.Fold = F2buf[cur];
        a2.mu = o.mu;
        a2.U = U1;
        a2.iters = F.it2.as<int>();
        a2.F = nullptr;
        a2.H21 = nullptr;
        a2.zncc = nullptr;
        a2.Ucopy = nullptr;
        a2.Fcopy = nullptr;
        a2.cc_mode = 2;
        CK(launch_icgn2(a2, ppl, st));
        if (relaunch) {
            IcgnArgs ar2 = a2;
            ar2.list = ar.list;
            ar2.nlist = ar.nlist;
            ar2.nlist_dev = ar.nlist_dev;
            CK(launch_icgn2(ar2, ppl_big, st));
        }
        // the step's count of statuses < 0 or > 99 (Subpb1.m:166-169) comes with the rule
        int *nbad_step = step < kMaxSteps ? F.nbad.as<int>() + step : nullptr;
        if (!o.no_subpb1_fit) {
            // Subpb1.m:165-199: bad nodes (status < 0 or > 99) carry no datum; gridfit of u
            // and v over the good nodes, written to every node.  The axis fill of the bad
            // nodes is only the CG's starting guess.
            CK(launch_badfill(F.it2.as<int>(), S, F.Nx, F.Ny, 0, nullptr, nullptr, U1, nullptr,
                              F.bad.as<int>(), F.part.as<double>(), st, nbad_step));
            CK(gfit(U1, 1, F.bad.as<int>()));
        } else {
            CK(launch_badmark(F.it2.as<int>(), S, 0, nullptr, nullptr, U1, nullptr,
                              F.bad.as<int>(), F.part.as<double>(), st, nbad_step));
            CK(natural_fill(F.nat, F.coords.as<double>(), F.bad.as<int>(), U1, nullptr, st,
                            F.nat_err.as<int>()));
        }
        F1p = F2buf[cur];  // FSubpb1 = FSubpb2 (main_ALDIC.m:364)
        cur ^= 1;
        if (stop) CK(hipMemcpyAsync(F.U1s.p, U1, 16 * (size_t)S, hipMemcpyDeviceToDevice, st));
        if (smooth && (r = smooth_field(U1, 2, F.Us.as<double>()))) return r;
        span(5, i0, smark());
        r = global_step(false, smooth);
        if (r) return r;
        ++F.admm_steps;
        if (stop) {
            double *errs = F.errs.as<double>();
            CK(launch_sqdiff(F.U2old.as<double>(), U2, 2 * (long)S, errs, st));
            CK(launch_sqdiff(F.U1old.as<double>(), F.U1s.as<double>(), 2 * (long)S, errs + 1, st));
            CK(hipMemcpyAsync(F.hostd, errs, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
            CK(hipMemcpyAsync(F.U2old.p, U2, 16 * (size_t)S, hipMemcpyDeviceToDevice, st));
            CK(hipMemcpyAsync(F.U1old.p, F.U1s.p, 16 * (size_t)S, hipMemcpyDeviceToDevice, st));
            CK(hipStreamSynchronize(st));
            F.upd = std::sqrt(F.hostd[0] / (2.0 * S));
            F.upd2 = std::sqrt(F.hostd[1] / (2.0 * S));
            if (F.upd < o.admm_tol || F.upd2 < o.admm_tol) break;
        }
    }
    F.f2cur = cur;
    F.gf_count = ngf;
    if (ngf > 0)
        CK(hipMemcpyAsync(F.gfh, F.gfstat.p, 3 * sizeof(double) * ngf, hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(F.nat_err_h, F.nat_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    const int e_end = mark();
    span(7, e_start, e_end);
    if (e_end < 0) return fail(ctx, ALDIC_ERR_HIP, "frame_run: event pool exhausted");
    CK(hipEventSynchronize(ev[e_end]));
    for (size_t k = 0; k + 2 < spans.size(); k += 3) {
        float t = 0;
        if (spans[k + 1] < 0 || spans[k + 2] < 0) continue;
        CK(hipEventElapsedTime(&t, ev[spans[k + 1]], ev[spans[k + 2]]));
        acc[spans[k]] += t;
    }
    for (int k = 0; k < 8; ++k) F.ms[k] = acc[k];
    // the search kernel alone (after its box-sum planes, if any) and the main 6-DOF kernel
    // alone (before the re-launch of enlarged windows)
    F.ms[8] = F.ms[9] = 0.f;
    if (e_search >= 0) CK(hipEventElapsedTime(&F.ms[8], F.ev_skern, ev[e_search]));
    if (e_search >= 0) CK(hipEventElapsedTime(&F.ms[9], ev[e_search], F.ev_ikern));
    if (*F.nat_err_h != 0)
        return fail(ctx, ALDIC_ERR_NUMERIC, "frame_run: a natural-neighbour cell could not be formed");
    return ALDIC_OK;
}

static int frame_run_impl(aldic_ctx *ctx, bool seeded)
{
    FrameState &F = ctx->fr;
    // The frame runs without a host round trip; should a gridfit have needed more than
    // its iteration budget, the frame runs again with polled gridfits (it depends only
    // on the staged images and options, so the second pass gives the polled result).
    int r = frame_pass(ctx, false, seeded);
    if (r) return r;
    if (F.gf_count == 0) return ALDIC_OK;
    bool ok = true;
    int most = 0;
    for (int k = 0; k < F.gf_count; ++k) {
        ok = ok && F.gfh[3 * k] != 0.0;
        most = std::max(most, (int)F.gfh[3 * k + 1]);
    }
    if (!ok) {
        ++F.gf_reruns;
        if ((r = frame_pass(ctx, true, seeded)) != ALDIC_OK) return r;
        for (int k = 0; k < F.gf_count; ++k) most = std::max(most, (int)F.gfh[3 * k + 1]);
    }
    // the next frame's budget: the most iterations seen here and a margin
    F.gf_most = most;
    F.gf_budget = std::min(64, std::max(6, most + 3));
    return ALDIC_OK;
}

int aldic_frame_run(aldic_ctx *ctx)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->fr.ready) return fail(ctx, ALDIC_ERR_STATE, "frame_run: frame_setup first");
    CK(hipSetDevice(ctx->device));
    return frame_run_impl(ctx, false);
}

int aldic_frame_run_seeded(aldic_ctx *ctx, const double *U0)
{
    if (!ctx) return ALDIC_ERR_ARG;
    FrameState &F = ctx->fr;
    if (!F.ready) return fail(ctx, ALDIC_ERR_STATE, "frame_run_seeded: frame_setup first");
    if (!U0) return fail(ctx, ALDIC_ERR_ARG, "frame_run_seeded: U0");
    CK(hipSetDevice(ctx->device));
    CK(up(F.seed, U0, 2 * (size_t)F.S, ctx->stream));
    return frame_run_impl(ctx, true);
}

int aldic_frame_scalars(aldic_ctx *ctx, double *out, int n)
{
    if (!ctx || !out || n < 0 || n > 4) return ALDIC_ERR_ARG;
    const FrameState &F = ctx->fr;
    const double v[4] = {F.beta_used, (double)F.admm_steps, F.upd, F.upd2};
    for (int k = 0; k < n; ++k) out[k] = v[k];
    return ALDIC_OK;
}

int aldic_qframe_setup(aldic_ctx *ctx, const aldic_qframe_opts *opts, const int *ws,
                       const int *tri, const double *w)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!ctx->have_quad) return fail(ctx, ALDIC_ERR_STATE, "qframe_setup: no quadtree mesh");
    if (!opts || !ws || !tri || !w || opts->n_admm < 0 || opts->cg_maxit < 1 ||
        !(opts->cg_rtol > 0) || !(opts->tol > 0) || !(opts->mu > 0) || !(opts->beta > 0))
        return fail(ctx, ALDIC_ERR_ARG, "qframe_setup: bad arguments");
    aldic::QuadDev &q = ctx->quad;
    QFrameState &Q = ctx->qf;
    const int S = q.m.nn, ngp = 4 * q.m.E;
    for (int k = 0; k < 3 * S; ++k)
        if (tri[k] < 0 || tri[k] >= ngp)
            return fail(ctx, ALDIC_ERR_ARG, "qframe_setup: Gauss point index");
    // 6-DOF launches: nodes grouped by pixel class in increasing node order
    std::vector<int> cls(S);
    for (int i = 0; i < S; ++i) {
        cls[i] = icgn_ppl_for(ws[i]);
        if (ws[i] < 2 || cls[i] == 0) return fail(ctx, ALDIC_ERR_ARG, "qframe_setup: window size");
    }
    std::vector<int> lists;
    Q.boff.clear(); Q.bn.clear(); Q.bppl.clear();
    for (int c : {5, 8, 18, 72}) {
        const int off = (int)lists.size();
        for (int i = 0; i < S; ++i)
            if (cls[i] == c) lists.push_back(i);
        if ((int)lists.size() > off) {
            Q.boff.push_back(off);
            Q.bn.push_back((int)lists.size() - off);
            Q.bppl.push_back(c);
        }
    }
    CK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t s8 = sizeof(double) * (size_t)S;
    CK(up(Q.ws, ws, S, st));
    CK(up(Q.lists, lists.data(), lists.size(), st));
    CK(up(Q.tri, tri, 3 * (size_t)S, st));
    CK(up(Q.w, w, 3 * (size_t)S, st));
    for (DBuf *b : {&Q.U0, &Q.U6, &Q.U1, &Q.U2, &Q.vdual, &Q.b}) CK(b->ensure(2 * s8));
    for (DBuf *b : {&Q.F6, &Q.F1, &Q.F2a, &Q.F2b, &Q.udual}) CK(b->ensure(4 * s8));
    CK(Q.H21.ensure(21 * s8));
    CK(Q.cc.ensure(8 * s8));
    CK(Q.Fgp.ensure(sizeof(double) * 4 * (size_t)(ngp ? ngp : 1)));
    CK(Q.it6.ensure(sizeof(int) * (size_t)S));
    CK(Q.it2.ensure(sizeof(int) * (size_t)S));
    CK(Q.bad.ensure(sizeof(int) * (size_t)S));
    CK(Q.known.ensure(sizeof(int) * 2 * (size_t)S));
    CK(Q.rl.ensure(sizeof(int) * ((size_t)S + 1 + (S + 4095) / 4096)));  // list, count, run counts
    CK(Q.part.ensure(sizeof(double) * 3 * 256));
    CK(Q.pstat.ensure(sizeof(double) * 3 * (size_t)(opts->n_admm + 1)));
    if (Q.hstat) (void)hipHostFree(Q.hstat);
    Q.hstat = nullptr;
    CK(hipHostMalloc((void **)&Q.hstat, sizeof(double) * 3 * (size_t)(opts->n_admm + 1)));
    CK(Q.nat_err.ensure(sizeof(int)));
    if (!Q.nat_err_h) CK(hipHostMalloc((void **)&Q.nat_err_h, sizeof(int), hipHostMallocDefault));
    for (auto &e : Q.ev)
        if (!e) CK(hipEventCreate(&e));
    CK(aldic::quad_assemble(q, opts->beta, opts->mu, st));
    CK(hipStreamSynchronize(st));
    Q.o = *opts;
    Q.snx = Q.sny = 0;
    if (opts->smooth_h > 0) {
        // funSmoothStrainQuadtree's grid min:h:max and each node's place on it
        if (!(opts->smoothness > 0)) return fail(ctx, ALDIC_ERR_ARG, "qframe_setup: smoothness must be > 0");
        std::vector<double> xy(2 * (size_t)S);
        CK(hipMemcpy(xy.data(), q.m.coords.p, sizeof(double) * 2 * (size_t)S, hipMemcpyDeviceToHost));
        std::vector<int> gidx;
        int nx = 0, ny = 0;
        if (!smooth_grid(S, xy.data(), opts->smooth_h, nx, ny, gidx))
            return fail(ctx, ALDIC_ERR_ARG, "qframe_setup: a node is off the smoothing grid min:smooth_h:max");
        CK(up(Q.sgidx, gidx.data(), gidx.size(), ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));   // gidx is a local
        Q.snx = nx;
        Q.sny = ny;
    }
    Q.S = S;
    Q.ready = true;
    return ALDIC_OK;
}

// One quadtree frame.  polled = false: every PCG gets pcg_budget iteration kernels and
// no host wait, its status goes to pstat; true: PCG polls every 8 iterations.
static int qframe_pass(aldic_ctx *ctx, const double *U0, bool polled)
{
    if (!ctx) return ALDIC_ERR_ARG;
    QFrameState &Q = ctx->qf;
    if (!Q.ready || !ctx->have_quad) return fail(ctx, ALDIC_ERR_STATE, "qframe_run: no setup");
    if (!ctx->have_images) return fail(ctx, ALDIC_ERR_STATE, "qframe_run: no images");
    if (!U0) return fail(ctx, ALDIC_ERR_ARG, "qframe_run: U0");
    aldic::QuadDev &q = ctx->quad;
    if (q.m.beta != Q.o.beta || q.m.mu != Q.o.mu) CK(aldic::quad_assemble(q, Q.o.beta, Q.o.mu, ctx->stream));
    CK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int S = Q.S;
    const aldic_qframe_opts &o = Q.o;
    CK(up(Q.U0, U0, 2 * (size_t)S, st));
    CK(hipEventRecord(Q.ev[0], st));
    CK(natural_err_clear(Q.nat_err.as<int>(), st));
    // LocalICGN: one launch per window class over its node list
    IcgnArgs a{};
    a.f = ctx->f.as<float>();
    a.g = ctx->g.as<float>();
    a.gx = ctx->gx.as<float>();
    a.gy = ctx->gy.as<float>();
    a.nx = ctx->nx;
    a.ny = ctx->ny;
    a.S = S;
    a.coords = q.m.coords.as<double>();
    a.ws = Q.ws.as<int>();
    a.tol = o.tol;
    a.U0 = Q.U0.as<double>();
    a.U = Q.U6.as<double>();
    a.F = Q.F6.as<double>();
    a.H21 = Q.H21.as<double>();
    a.iters = Q.it6.as<int>();
    a.Ucopy = Q.U1.as<double>();  // bad-point step works on the copies
    a.Fcopy = Q.F1.as<double>();
    a.cc = Q.cc.as<double>();     // window constants kept for the 2-DOF runs
    a.cc_mode = 1;
    for (size_t k = 0; k < Q.bn.size(); ++k) {
        IcgnArgs ak = a;
        ak.list = Q.lists.as<int>() + Q.boff[k];
        ak.nlist = Q.bn[k];
        CK(launch_icgn6(ak, Q.bppl[k], st));
    }
    // windows the mask rule enlarged past their class (status -2) run again in the
    // largest class
    {
        IcgnArgs ar = a;
        int *rl = Q.rl.as<int>();
        CK(launch_list_eq(Q.it6.as<int>(), S, -2, rl, rl + S, rl + S + 1, st));
        ar.list = rl;
        ar.nlist = S;
        ar.nlist_dev = rl + S;
        CK(launch_icgn6(ar, 72, st));
    }
    double *U1 = Q.U1.as<double>(), *F1 = Q.F1.as<double>(), *U2 = Q.U2.as<double>();
    CK(launch_badmark(Q.it6.as<int>(), S, 1, nullptr, nullptr, U1, F1, Q.bad.as<int>(),
                      Q.part.as<double>(), st));
    CK(natural_fill(q.nat, q.m.coords.as<double>(), Q.bad.as<int>(), U1, F1, st,
                    Q.nat_err.as<int>()));
    CK(hipEventRecord(Q.ev[1], st));
    // funSmoothStrainQuadtree (main_ALDIC_Quadtree.m:170-172, 221, 283-284), when on
    const bool smooth = o.smooth_h > 0 && Q.snx > 0;
    Q.smooth_iters = 0;
    auto smooth_F = [&](double *Fs) -> hipError_t {
        int it[2] = {0, 0};
        const hipError_t e = smooth_strain_qt(S, Q.sgidx.as<int>(), Q.snx, Q.sny, o.smoothness, 0.5, 1e-12, 20000,
                                              Fs, Q.sw, it, st);
        Q.smooth_iters += it[0] + it[1];
        return e;
    };
    if (smooth) CK(smooth_F(F1));

    double *F2buf[2] = {Q.F2a.as<double>(), Q.F2b.as<double>()};
    int cur = 0;
    const double *F1p = F1;
    static thread_local DBuf work[8];
    int nglob = 0;
    float t_glob = 0.f;
    auto global_step = [&](bool first) -> int {
        CK(hipEventRecord(Q.ev[2], st));
        CK(aldic::quad_rhs(q, o.beta, o.mu, U1, F1p, Q.b.as<double>(), st));
        double *rec = Q.pstat.as<double>() + 3 * nglob++;
        // warm start from Subpb1's U (the global step moves it little)
        CK(aldic::quad_start(q, U1, U2, st));
        CK(aldic::fem_pcg(q.m, Q.b.as<double>(), U2, o.cg_rtol,
                          polled ? o.cg_maxit : std::min(o.cg_maxit, Q.pcg_budget), work, nullptr,
                          st, polled ? 8 : 0, rec, true));
        CK(aldic::quad_finish(q, U1, U2, st));
        CK(aldic::quad_gp_strain(q, U2, Q.Fgp.as<double>(), nullptr, st));
        CK(aldic::quad_gp_gather(S, Q.tri.as<int>(), Q.w.as<double>(), Q.Fgp.as<double>(),
                                 F2buf[cur], st));
        // hole-edge nodes keep Subpb1's F (main_ALDIC_Quadtree.m:222,225,286,288)
        CK(aldic::quad_copy_dirichlet(q, F1p, F2buf[cur], st));
        if (smooth) {
            // F2 = smooth(0.1 F2 + 0.9 F1), the hole edge copied again in the ADMM loop
            CK(launch_blend_f(F2buf[cur], F1p, 4 * (long)S, st));
            CK(smooth_F(F2buf[cur]));
            if (!first) CK(aldic::quad_copy_dirichlet(q, F1p, F2buf[cur], st));
        }
        CK(launch_sub(F2buf[cur], F1p, Q.udual.as<double>(), 4 * (long)S, st));
        CK(launch_sub(U2, U1, Q.vdual.as<double>(), 2 * (long)S, st));
        CK(hipEventRecord(Q.ev[3], st));
        if (polled) {
            CK(hipEventSynchronize(Q.ev[3]));
            float ms = 0.f;
            CK(hipEventElapsedTime(&ms, Q.ev[2], Q.ev[3]));
            t_glob += ms;
        }
        return ALDIC_OK;
    };
    int r = global_step(true);
    if (r) return r;
    for (int step = 0; step < o.n_admm; ++step) {
        // Subpb1Quadtree from (U2, F2) with the LocalICGN Hessian, on the same windows
        IcgnArgs a2 = a;
        a2.U0 = nullptr;
        a2.Hin = Q.H21.as<double>();
        a2.vdual = Q.vdual.as<double>();
        a2.Uold = U2;
        a2.Fold = F2buf[cur];
        a2.mu = o.mu;
        a2.U = U1;
        a2.iters = Q.it2.as<int>();
        a2.F = nullptr;
        a2.H21 = nullptr;
        a2.Ucopy = nullptr;
        a2.Fcopy = nullptr;
        a2.cc_mode = 2;
        for (size_t k = 0; k < Q.bn.size(); ++k) {
            IcgnArgs ak = a2;
            ak.list = Q.lists.as<int>() + Q.boff[k];
            ak.nlist = Q.bn[k];
            CK(launch_icgn2(ak, Q.bppl[k], st));
        }
        {
            IcgnArgs ar2 = a2;
            int *rl = Q.rl.as<int>();
            CK(launch_list_eq(Q.it2.as<int>(), S, -2, rl, rl + S, rl + S + 1, st));
            ar2.list = rl;
            ar2.nlist = S;
            ar2.nlist_dev = rl + S;
            CK(launch_icgn2(ar2, 72, st));
        }
        CK(launch_badmark(Q.it2.as<int>(), S, 2, nullptr, nullptr, U1, nullptr, Q.bad.as<int>(),
                          Q.part.as<double>(), st));
        CK(natural_fill(q.nat, q.m.coords.as<double>(), Q.bad.as<int>(), U1, nullptr, st,
                        Q.nat_err.as<int>()));
        F1p = F2buf[cur];  // FSubpb1 = FSubpb2 (main_ALDIC_Quadtree.m:258)
        cur ^= 1;
        if ((r = global_step(false))) return r;
    }
    Q.cur_F2 = cur;
    // one wait per frame: the PCG records
    CK(hipMemcpyAsync(Q.hstat, Q.pstat.p, 3 * sizeof(double) * nglob, hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(Q.nat_err_h, Q.nat_err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    CK(hipEventSynchronize(Q.ev[3]));
    CK(hipStreamSynchronize(st));
    float ms_frame = 0.f, ms_local = 0.f;
    CK(hipEventElapsedTime(&ms_frame, Q.ev[0], Q.ev[3]));
    CK(hipEventElapsedTime(&ms_local, Q.ev[0], Q.ev[1]));
    Q.t_frame = ms_frame;
    Q.t_local = ms_local;
    Q.t_global = polled ? t_glob : NAN;
    int total = 0, most = 0;
    bool ok = true;
    for (int k = 0; k < nglob; ++k) {
        // converged: the stop flag AND a finite relative residual within the tolerance (the
        // flag is also set on a breakdown or a NaN residual)
        const double rel = Q.hstat[3 * k + 2];
        ok = ok && Q.hstat[3 * k] != 0.0 && std::isfinite(rel) && rel <= Q.o.cg_rtol;
        total += (int)Q.hstat[3 * k + 1];
        most = std::max(most, (int)Q.hstat[3 * k + 1]);
    }
    Q.pcg_iters = total;
    Q.pcg_most = most;
    if (*Q.nat_err_h != 0)
        return fail(ctx, ALDIC_ERR_NUMERIC, "qframe_run: a natural-neighbour cell could not be formed");
    return ok ? ALDIC_OK : -1000;  // a PCG ran out of its budget
}

int aldic_qframe_run(aldic_ctx *ctx, const double *U0)
{
    if (!ctx) return ALDIC_ERR_ARG;
    QFrameState &Q = ctx->qf;
    int r = qframe_pass(ctx, U0, false);
    if (r == -1000) {
        // a global step needed more than the budget: run the frame again with polled PCGs
        // and size the budget for the next frames from its iteration counts
        ++Q.pcg_reruns;
        r = qframe_pass(ctx, U0, true);
        if (r == -1000) return fail(ctx, ALDIC_ERR_NUMERIC, "qframe_run: PCG did not converge");
    }
    if (r == ALDIC_OK) Q.pcg_budget = std::min(Q.o.cg_maxit, std::max(12, Q.pcg_most + 4));
    return r;
}

int aldic_qframe_get(aldic_ctx *ctx, double *U, double *F, int *iters6, int *iters2,
                     double *times)
{
    if (!ctx) return ALDIC_ERR_ARG;
    QFrameState &Q = ctx->qf;
    if (!Q.ready) return fail(ctx, ALDIC_ERR_STATE, "qframe_get: no setup");
    CK(hipSetDevice(ctx->device));
    const size_t S = Q.S;
    if (U) CK(down(U, Q.U2.p, 2 * S, ctx->stream));
    if (F) CK(down(F, (Q.cur_F2 ? Q.F2b : Q.F2a).p, 4 * S, ctx->stream));
    if (iters6) CK(down(iters6, Q.it6.p, S, ctx->stream));
    if (iters2) CK(down(iters2, Q.it2.p, S, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (times) {
        times[0] = Q.t_frame;
        times[1] = Q.t_local;
        times[2] = Q.t_global;
        times[3] = Q.pcg_iters;
        times[4] = Q.pcg_budget;
        times[5] = Q.pcg_reruns;
    }
    return ALDIC_OK;
}

int aldic_frame_counters(aldic_ctx *ctx, int *out, int n)
{
    if (!ctx || !out || n < 0 || n > 4 + kMaxSteps) return ALDIC_ERR_ARG;
    int v[4 + kMaxSteps] = {ctx->fr.gf_budget, ctx->fr.gf_reruns, ctx->fr.gf_most};
    for (int k = 0; k < kMaxSteps; ++k) v[3 + k] = -1;
    v[3 + kMaxSteps] = ctx->fr.n_fallback;
    if (n > 3 && ctx->fr.ready && ctx->fr.nbad.p) {
        CK(hipSetDevice(ctx->device));
        CK(hipStreamSynchronize(ctx->stream));
        CK(hipMemcpy(v + 3, ctx->fr.nbad.p, sizeof(int) * kMaxSteps, hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < n; ++k) out[k] = v[k];
    return ALDIC_OK;
}

int aldic_frame_get(aldic_ctx *ctx, double *U, double *Fo, double *U6, double *F6, int *iters6,
                    int *iters2, double *u0)
{
    if (!ctx) return ALDIC_ERR_ARG;
    FrameState &F = ctx->fr;
    if (!F.ready) return fail(ctx, ALDIC_ERR_STATE, "frame_get: no frame");
    CK(hipSetDevice(ctx->device));
    const size_t S = F.S;
    CK(down(U, F.U2.p, 2 * S, ctx->stream));
    CK(down(Fo, F.f2cur ? F.F2b.p : F.F2.p, 4 * S, ctx->stream));
    CK(down(U6, F.U6.p, 2 * S, ctx->stream));
    CK(down(F6, F.F6.p, 4 * S, ctx->stream));
    CK(down(iters6, F.it6.p, S, ctx->stream));
    CK(down(iters2, F.it2.p, S, ctx->stream));
    CK(down(u0, F.U0.p, 2 * S, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    return ALDIC_OK;
}

int aldic_frame_times(aldic_ctx *ctx, double *ms, int n)
{
    if (!ctx || !ms || n < 0 || n > 10) return ALDIC_ERR_ARG;
    for (int k = 0; k < n; ++k) ms[k] = ctx->fr.ms[k];
    return ALDIC_OK;
}

}  // extern "C"
