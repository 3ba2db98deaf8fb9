/*
 * natural_oracle.c -- FP64 CPU restatement of the bad-node refill of LocalICGN and
 * Subpb1Quadtree: scatteredInterpolant(good xy, good values, 'natural') evaluated at
 * every node (LocalICGN.m:141-161, Subpb1Quadtree.m:188-199; MATLAB's default
 * ExtrapolationMethod for 'natural' is 'linear').
 *
 * TEST INFRASTRUCTURE ONLY: the tests and the CPU-baseline leg of the benchmark load it
 * as the oracle the HIP library is compared against.  Nothing in the product links it.
 *
 * Definitions restated here (the interpolant's, not MATLAB's code, which is closed):
 *   - good nodes keep their values (the interpolant passes through its data);
 *   - a bad node q strictly inside the convex hull of the good nodes P gets Sibson's
 *     natural-neighbour interpolant: with V(q) the Voronoi cell of q in the diagram of
 *     P + {q} and V_P(p) the cell of p in the diagram of P, the weight of p is
 *     area(V(q) & V_P(p)) / area(V(q)).  V(q) is formed by clipping a large square
 *     around q with the half-planes {x : |x - q| <= |x - p|}, p in P, in order of
 *     distance, until the next p is at least twice the cell's radius away (such a p
 *     cannot cut the cell); the part of V(q) nearest p_i is V(q) clipped with
 *     {x : |x - p_i| <= |x - p_j|} for the other natural neighbours p_j (inside V(q)
 *     only natural neighbours' cells occur).  Every polygon is kept as its cyclic list
 *     of lines and every vertex is the intersection of two of them, so the vertices are
 *     circumcentres of site triples computed from the sites themselves;
 *   - a bad node on the hull boundary or outside it (parity unpinned: MATLAB's 'linear'
 *     extrapolation is documented only as "based on boundary gradients"; this is the
 *     rule chosen, linear-exact like Sibson's interpolant): let b be the point of the
 *     hull nearest q.  If b is inside a hull edge, a and c are the good nodes on that
 *     edge next to b (a at or before b, c after it, counter-clockwise) and r is the
 *     Delaunay third vertex of the edge a-c (the good node on the inner side whose
 *     circle through a and c bulges least inward; ties -> lowest node index); q takes
 *     the value of the plane through a, c and r.  If b is a hull vertex v, q takes the
 *     plane through v and the good nodes next to v along its two hull edges.  On the
 *     boundary (b = q) both reduce to linear interpolation along the hull edge, the
 *     limit of the natural-neighbour interpolant there;
 *   - fewer than three non-collinear good nodes: the nearest good node's value (ties ->
 *     lowest index); no good node at all: NaN.
 *
 * Pinning (tests/test_oracle_natural.py): an isolated hole on a square lattice gets
 * exactly 1/4 from each edge neighbour; weights are non-negative, sum to 1 and reproduce
 * the node's position (linear precision), so linear fields come back exactly around
 * single, paired, L-shaped, 3 x 3 and random clusters and on the hull; brute-force
 * area sampling of the nearest-site map on small random and lattice configurations.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NN_MAXL 512

typedef struct {
    double nx, ny, c; /* half-plane nx*x + ny*y <= c (coordinates relative to q) */
    int tag;          /* good-node index, or -1 for the initial square */
} nline;

typedef struct {
    int m;
    nline L[NN_MAXL];
    double vx[NN_MAXL], vy[NN_MAXL]; /* vertex k = L[k] meet L[k+1] */
} npoly;

static void nn_meet(const nline *a, const nline *b, double *x, double *y)
{
    const double det = a->nx * b->ny - a->ny * b->nx;
    *x = (a->c * b->ny - b->c * a->ny) / det;
    *y = (a->nx * b->c - b->nx * a->c) / det;
}

static void nn_vertices(npoly *P)
{
    for (int k = 0; k < P->m; ++k) nn_meet(&P->L[k], &P->L[(k + 1) % P->m], &P->vx[k], &P->vy[k]);
}

/* Clip P with M (keep M.n . x <= M.c).  0 ok, -1 overflow, -2 empty. */
static int nn_clip(npoly *P, const nline *M)
{
    const int m = P->m;
    static __thread int out[NN_MAXL];
    int nout = 0;
    for (int k = 0; k < m; ++k) {
        const double a = M->nx * P->vx[k], b = M->ny * P->vy[k];
        const double d = a + b - M->c;
        const double tol = 1e-12 * (fabs(a) + fabs(b) + fabs(M->c));
        out[k] = d > tol;
        nout += out[k];
    }
    if (nout == 0) return 0;
    if (nout == m) { P->m = 0; return -2; }
    int a = 0;
    while (!(out[a] && !out[(a - 1 + m) % m])) ++a;
    int len = 0;
    while (out[(a + len) % m]) ++len;
    const int nm = m - len + 2;
    if (nm > NN_MAXL) return -1;
    nline tmp[NN_MAXL];
    int j = 0;
    for (int k = 0; k < m - len + 1; ++k) tmp[j++] = P->L[(a + len + k) % m];  /* L[b+1] .. L[a] */
    tmp[j++] = *M;
    memcpy(P->L, tmp, sizeof(nline) * nm);
    P->m = nm;
    nn_vertices(P);
    return 0;
}

static double nn_area(const npoly *P)
{
    double s = 0.0;
    for (int k = 0; k < P->m; ++k) {
        const int k0 = (k + P->m - 1) % P->m;
        s += P->vx[k0] * P->vy[k] - P->vx[k] * P->vy[k0];
    }
    return 0.5 * s;
}

static double nn_radius(const npoly *P)
{
    double r = 0.0;
    for (int k = 0; k < P->m; ++k) {
        const double d = sqrt(P->vx[k] * P->vx[k] + P->vy[k] * P->vy[k]);
        if (d > r) r = d;
    }
    return r;
}

typedef struct { double d; int i; } nn_key;

static int nn_key_cmp(const void *a, const void *b)
{
    const nn_key *x = a, *y = b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->i > y->i) - (x->i < y->i);
}

/* Sibson coordinates of q w.r.t. the np points pxy (q strictly inside their hull).
 * idx / w receive the natural neighbours (indices into pxy) and their weights.
 * Returns their number, or -1 (cell unbounded or too many lines), -2 (maxn too small). */
int orc_sibson(int np, const double *pxy, double qx, double qy, int *idx, double *w, int maxn)
{
    if (np < 3) return -1;
    nn_key *key = malloc(sizeof(nn_key) * (size_t)np);
    double ext = 0.0;
    for (int i = 0; i < np; ++i) {
        const double dx = pxy[2 * i] - qx, dy = pxy[2 * i + 1] - qy;
        key[i].d = sqrt(dx * dx + dy * dy);
        key[i].i = i;
        if (fabs(dx) > ext) ext = fabs(dx);
        if (fabs(dy) > ext) ext = fabs(dy);
    }
    qsort(key, (size_t)np, sizeof(nn_key), nn_key_cmp);
    npoly *V = malloc(sizeof(npoly)), *R = malloc(sizeof(npoly));
    const double B = 1e6 * (ext + 1.0);
    const double bn[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
    V->m = 4;
    for (int k = 0; k < 4; ++k) {
        V->L[k].nx = bn[k][0]; V->L[k].ny = bn[k][1]; V->L[k].c = B; V->L[k].tag = -1;
    }
    nn_vertices(V);
    double rho = nn_radius(V);
    int rc = 0;
    for (int s = 0; s < np; ++s) {
        if (key[s].d >= 2.0 * rho) break;
        const int i = key[s].i;
        const double px = pxy[2 * i] - qx, py = pxy[2 * i + 1] - qy;
        nline M = {px, py, 0.5 * (px * px + py * py), i};
        if (nn_clip(V, &M)) { rc = -1; break; }
        rho = nn_radius(V);
    }
    for (int k = 0; rc == 0 && k < V->m; ++k)
        if (V->L[k].tag < 0) rc = -1;
    if (rc == 0 && V->m > maxn) rc = -2;
    if (rc == 0) {
        const int nnb = V->m;
        double tot = 0.0;
        for (int k = 0; k < nnb; ++k) {
            const int i = V->L[k].tag;
            const double pix = pxy[2 * i] - qx, piy = pxy[2 * i + 1] - qy;
            memcpy(R, V, sizeof(npoly));
            int ok = 1;
            for (int l = 0; l < nnb && ok; ++l) {
                if (l == k) continue;
                const int j = V->L[l].tag;
                const double pjx = pxy[2 * j] - qx, pjy = pxy[2 * j + 1] - qy;
                nline M = {pjx - pix, pjy - piy,
                           0.5 * ((pjx * pjx + pjy * pjy) - (pix * pix + piy * piy)), j};
                const int c = nn_clip(R, &M);
                if (c == -2) ok = 0;         /* nothing of V(q) is nearest p_i */
                else if (c == -1) { rc = -1; break; }
            }
            if (rc) break;
            idx[k] = i;
            w[k] = ok ? nn_area(R) : 0.0;
            tot += w[k];
        }
        if (rc == 0) {
            for (int k = 0; k < nnb; ++k) w[k] /= tot;
            rc = nnb;
        }
    }
    free(V); free(R); free(key);
    return rc;
}

static double cross3(double ax, double ay, double bx, double by, double cx, double cy)
{
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
}

typedef struct { double x, y; int i; } nn_pt;

static int nn_pt_cmp(const void *a, const void *b)
{
    const nn_pt *p = a, *q = b;
    if (p->x != q->x) return p->x < q->x ? -1 : 1;
    if (p->y != q->y) return p->y < q->y ? -1 : 1;
    return (p->i > q->i) - (p->i < q->i);
}

/* Strict convex hull (collinear points dropped), counter-clockwise, Andrew's monotone
 * chain.  hull receives point indices; returns the vertex count. */
static int nn_hull(int n, const nn_pt *pts_sorted, int *hull)
{
    int *H = malloc(sizeof(int) * (size_t)(2 * n + 2));
    int k = 0;
    for (int s = 0; s < n; ++s) {
        while (k >= 2 && cross3(pts_sorted[H[k - 2]].x, pts_sorted[H[k - 2]].y,
                                pts_sorted[H[k - 1]].x, pts_sorted[H[k - 1]].y,
                                pts_sorted[s].x, pts_sorted[s].y) <= 0.0)
            --k;
        H[k++] = s;
    }
    for (int s = n - 2, t = k + 1; s >= 0; --s) {
        while (k >= t && cross3(pts_sorted[H[k - 2]].x, pts_sorted[H[k - 2]].y,
                                pts_sorted[H[k - 1]].x, pts_sorted[H[k - 1]].y,
                                pts_sorted[s].x, pts_sorted[s].y) <= 0.0)
            --k;
        H[k++] = s;
    }
    const int h = k > 1 ? k - 1 : k;
    for (int s = 0; s < h; ++s) hull[s] = pts_sorted[H[s]].i;
    free(H);
    return h;
}

/* Plane through nodes a, b, c (values va, vb, vc, nc components) evaluated at (x, y). */
static void nn_plane(const double *xy, int a, int b, int c, double x, double y, const double *V,
                     int nc, double *out)
{
    const double ax = xy[2 * a], ay = xy[2 * a + 1];
    const double bx = xy[2 * b] - ax, by = xy[2 * b + 1] - ay;
    const double cx = xy[2 * c] - ax, cy = xy[2 * c + 1] - ay;
    const double qx = x - ax, qy = y - ay;
    const double den = bx * cy - by * cx;
    const double al = (qx * cy - qy * cx) / den, be = (bx * qy - by * qx) / den;
    for (int k = 0; k < nc; ++k)
        out[k] = V[(size_t)nc * a + k] + al * (V[(size_t)nc * b + k] - V[(size_t)nc * a + k])
               + be * (V[(size_t)nc * c + k] - V[(size_t)nc * a + k]);
}

/* On segment A->C (within the relative tolerance), parameter s in [0, 1]. */
static int nn_on_seg(const double *xy, int A, int C, int p, double *s)
{
    const double ax = xy[2 * A], ay = xy[2 * A + 1];
    const double ex = xy[2 * C] - ax, ey = xy[2 * C + 1] - ay;
    const double px = xy[2 * p] - ax, py = xy[2 * p + 1] - ay;
    const double L2 = ex * ex + ey * ey;
    const double cr = ex * py - ey * px;
    if (fabs(cr) > 1e-12 * sqrt(L2) * sqrt(px * px + py * py)) return 0;
    *s = (px * ex + py * ey) / L2;
    return *s >= 0.0 && *s <= 1.0;
}

/* Circle through a and c (edge vector e, length L, midpoint m, inner unit normal n) and
 * node p on the inner side: its centre sits at m + s n; returns 0 for p not strictly
 * on the inner side. */
static int nn_bulge(const double *xy, double ax, double ay, double ex, double ey, double L,
                    double mx, double my, double nxv, double nyv, int p, double *s)
{
    const double px = xy[2 * p] - ax, py = xy[2 * p + 1] - ay;
    const double cr = ex * py - ey * px;
    if (!(cr > 1e-12 * L * sqrt(px * px + py * py))) return 0;
    const double dx = xy[2 * p] - mx, dy = xy[2 * p + 1] - my;
    *s = (dx * dx + dy * dy - 0.25 * L * L) / (2.0 * (dx * nxv + dy * nyv));
    return 1;
}

/* Fills bad nodes in place (nc values per node, V[nc*i + k]); kind (nullable) per node:
 * 0 good, 1 Sibson, 2 hull edge, 3 hull vertex, 4 nearest good node, 5 no good node.
 * Returns 0, or -1 when a Sibson cell could not be formed. */
int orc_natural_fill(int nn, const double *xy, const int *bad, int nc, double *V, int *kind)
{
    int ng = 0;
    for (int i = 0; i < nn; ++i) ng += !bad[i];
    int *gid = malloc(sizeof(int) * (size_t)(ng + 1));
    double *gxy = malloc(sizeof(double) * 2 * (size_t)(ng + 1));
    nn_pt *pts = malloc(sizeof(nn_pt) * (size_t)(ng + 1));
    int *hull = malloc(sizeof(int) * (size_t)(2 * ng + 2));
    for (int i = 0, j = 0; i < nn; ++i)
        if (!bad[i]) {
            gid[j] = i;
            gxy[2 * j] = xy[2 * i]; gxy[2 * j + 1] = xy[2 * i + 1];
            pts[j].x = xy[2 * i]; pts[j].y = xy[2 * i + 1]; pts[j].i = i;
            ++j;
        }
    qsort(pts, (size_t)ng, sizeof(nn_pt), nn_pt_cmp);
    const int h = ng >= 3 ? nn_hull(ng, pts, hull) : 0;
    int err = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
    for (int q = 0; q < nn; ++q) {
        if (!bad[q]) { if (kind) kind[q] = 0; continue; }
        const double qx = xy[2 * q], qy = xy[2 * q + 1];
        double *out = V + (size_t)nc * q;
        if (ng == 0) {
            for (int k = 0; k < nc; ++k) out[k] = NAN;
            if (kind) kind[q] = 5;
            continue;
        }
        if (h < 3) {
            int best = -1;
            double bd = 0.0;
            for (int j = 0; j < ng; ++j) {
                const double dx = gxy[2 * j] - qx, dy = gxy[2 * j + 1] - qy, d = dx * dx + dy * dy;
                if (best < 0 || d < bd) { best = gid[j]; bd = d; }
            }
            for (int k = 0; k < nc; ++k) out[k] = V[(size_t)nc * best + k];
            if (kind) kind[q] = 4;
            continue;
        }
        int inside = 1;
        for (int e = 0; e < h && inside; ++e) {
            const int A = hull[e], C = hull[(e + 1) % h];
            const double ex = xy[2 * C] - xy[2 * A], ey = xy[2 * C + 1] - xy[2 * A + 1];
            const double px = qx - xy[2 * A], py = qy - xy[2 * A + 1];
            const double cr = ex * py - ey * px;
            if (!(cr > 1e-12 * sqrt(ex * ex + ey * ey) * sqrt(px * px + py * py))) inside = 0;
        }
        if (inside) {
            int idx[NN_MAXL];
            double w[NN_MAXL];
            const int n = orc_sibson(ng, gxy, qx, qy, idx, w, NN_MAXL);
            if (n < 0) {
                err |= 1;
                for (int k = 0; k < nc; ++k) out[k] = NAN;
                continue;
            }
            for (int k = 0; k < nc; ++k) {
                double s = 0.0;
                for (int t = 0; t < n; ++t) s += w[t] * V[(size_t)nc * gid[idx[t]] + k];
                out[k] = s;
            }
            if (kind) kind[q] = 1;
            continue;
        }
        /* nearest point of the hull */
        int be = 0;
        double bt = 0.0, bd = INFINITY;
        for (int e = 0; e < h; ++e) {
            const int A = hull[e], C = hull[(e + 1) % h];
            const double ax = xy[2 * A], ay = xy[2 * A + 1];
            const double ex = xy[2 * C] - ax, ey = xy[2 * C + 1] - ay;
            double t = ((qx - ax) * ex + (qy - ay) * ey) / (ex * ex + ey * ey);
            t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
            const double dx = ax + t * ex - qx, dy = ay + t * ey - qy, d = dx * dx + dy * dy;
            if (d < bd) { bd = d; be = e; bt = t; }
        }
        double res[8];
        if (bt > 0.0 && bt < 1.0) {
            const int A = hull[be], C = hull[(be + 1) % h];
            int a = A, c = C;
            double sa = 0.0, sc = 1.0;
            for (int j = 0; j < ng; ++j) {
                double s;
                if (!nn_on_seg(xy, A, C, gid[j], &s)) continue;
                if (s <= bt && s > sa) { sa = s; a = gid[j]; }
                if (s > bt && s < sc) { sc = s; c = gid[j]; }
            }
            /* Delaunay third vertex of a-c on the inner (left) side */
            const double ax = xy[2 * a], ay = xy[2 * a + 1];
            const double ex = xy[2 * c] - ax, ey = xy[2 * c + 1] - ay;
            const double L = sqrt(ex * ex + ey * ey);
            const double mx = ax + 0.5 * ex, my = ay + 0.5 * ey, nxv = -ey / L, nyv = ex / L;
            /* smallest bulge s, then the lowest node index among those within the
             * tolerance of it (co-circular candidates) */
            double rs = INFINITY;
            for (int j = 0; j < ng; ++j) {
                double s;
                if (nn_bulge(xy, ax, ay, ex, ey, L, mx, my, nxv, nyv, gid[j], &s) && s < rs) rs = s;
            }
            int r = -1;
            for (int j = 0; j < ng && r < 0; ++j) {
                double s;
                if (nn_bulge(xy, ax, ay, ex, ey, L, mx, my, nxv, nyv, gid[j], &s) &&
                    s <= rs + 1e-12 * (fabs(rs) + L))
                    r = gid[j];
            }
            nn_plane(xy, a, c, r, qx, qy, V, nc, res);
            if (kind) kind[q] = 2;
        } else {
            const int vi = bt <= 0.0 ? be : (be + 1) % h;
            const int v = hull[vi], P = hull[(vi - 1 + h) % h], Nn = hull[(vi + 1) % h];
            int u1 = P, u2 = Nn;
            double s1 = 0.0, s2 = 1.0;
            for (int j = 0; j < ng; ++j) {
                const int p = gid[j];
                double s;
                if (p == v) continue;
                if (nn_on_seg(xy, P, v, p, &s) && s < 1.0 && s > s1) { s1 = s; u1 = p; }
                if (nn_on_seg(xy, v, Nn, p, &s) && s > 0.0 && s < s2) { s2 = s; u2 = p; }
            }
            nn_plane(xy, v, u2, u1, qx, qy, V, nc, res);
            if (kind) kind[q] = 3;
        }
        for (int k = 0; k < nc; ++k) out[k] = res[k];
    }
    free(gid); free(gxy); free(pts); free(hull);
    return err ? -1 : 0;
}
