// Bad-node refill of LocalICGN and Subpb1Quadtree: scatteredInterpolant(good xy, good
// values, 'natural') evaluated at every node (LocalICGN.m:141-161,
// Subpb1Quadtree.m:188-199; MATLAB's default extrapolation for 'natural' is 'linear').
//
//   Good nodes keep their values.  A bad node q strictly inside the convex hull of the
//   good nodes P takes Sibson's natural-neighbour interpolant: the weight of p is the
//   area the Voronoi cell V(q) of q (in the diagram of P + {q}) takes from p's cell, over
//   area(V(q)).  V(q) = a large square clipped by the half-planes |x - q| <= |x - p|,
//   p taken from rings of buckets around q until the next ring is at least twice the
//   cell's radius away; the part of V(q) nearest p_i = V(q) clipped by
//   |x - p_i| <= |x - p_j| over the other natural neighbours p_j.  Polygons are cyclic
//   lists of lines (each entry: the line's tag and the vertex where it meets the next
//   line), every vertex the meet of two lines given by the sites themselves.
//   A bad node on or outside the hull (the rule chosen for MATLAB's 'linear'
//   extrapolation, which is documented only as "based on boundary gradients"): b = the
//   nearest hull point; inside a hull edge, the plane through the good nodes a, c next
//   to b on that edge and the edge's Delaunay third vertex r (least inward bulge, ties to
//   the lowest node index); at a hull vertex v, the plane through v and its neighbours
//   along the two hull edges.  Fewer than three non-collinear good nodes: the nearest
//   good node (ties to the lowest index); none: NaN.
//
// Kernels (one refill = 7 launches, no host wait):
//   k_nat_rows      one wave per row of equal y: first / last good node, good count
//   k_nat_hull      one workgroup: good count, the row extremes pruned by their
//                   eight-direction extreme polygon, strict convex hull of the rest
//                   (monotone chain), counters cleared
//   k_nat_classify  per bad node: inside the hull -> inside list, else -> hull list
//   k_nat_sibson    one thread per inside node, its two polygons in LDS (24 lines);
//                   a cell with more natural neighbours, or still open past ring
//                   kNatRingCap, goes to a list for
//   k_nat_sibson_wave one wave per node: the cell built by wave-parallel clips in LDS
//                   (512 lines), each natural neighbour's region on its own lane;
//                   nodes beyond those limits go to
//   k_nat_sibson_big one thread per node, polygons of up to 512 lines in device memory
//   k_nat_hullfill  one wave per hull-list node, scans over rings of buckets
#include <algorithm>
#include <cmath>
#include <vector>

#include "aldic_internal.h"
#include "common.h"

namespace aldic {

namespace {

constexpr int kNatMax = 24;        // lines per polygon in LDS (first pass)
constexpr int kNatThreads = 64;    // k_nat_sibson workgroup
constexpr int kNatBig = 512;       // lines per polygon in the second pass (device memory)
constexpr int kNatBigThreads = 1024;  // second-pass threads (polygon slots)
constexpr int kRegion = 0x40000000;  // tag bit: bisector of p_i and p_(tag & ~kRegion)
constexpr int kNatRingCap = 1;     // first pass: a cell still open past ring 1 -> big list
constexpr int kNatWaveV = 128;     // k_nat_sibson_wave: lines of the cell (LDS, two buffers)
constexpr int kNatWaveR = 24;      // lines of each lane's region polygon (LDS)
constexpr int kNatWaveCand = 1536; // candidates gathered per chunk of 64 ring buckets
constexpr int kNatPre = 4;         // nodes per ring bucket read with their loads in flight
constexpr int kNatLaneB = 4;       // ring buckets per lane in a chunk of the wave pass
constexpr int kNatAllWave = 8192;  // up to this many inside bad nodes, all go to the wave pass

struct NatArgs {
    const double2 *xy;
    const int *bad;
    double2 *U;      // 2 values per node, in place
    double4 *F;      // 4 values per node, in place (nullable)
    int nn;
    // buckets
    int BX, BY;
    double ox, oy, h;
    const int *bstart, *bnode;
    // rows (equal y, x ascending)
    int R;
    const int *rstart, *rnode;
    // work
    int *rows;       // 3 R: first, last, count
    int *chain;      // 6 R + 4: row extremes, then the chain stack
    int *hull;       // 2 R + 2 (counter-clockwise, from the lowest-x, lowest-y vertex)
    int *ctl;        // {hull size, good count, n inside, n hull, errors, n big}
    int *lin, *lhull, *lbig;  // nn each
    int *lbig2;      // nn: nodes beyond k_nat_sibson_wave's limits (ctl[6] of them)
    int *err;        // nullable: refills whose cell could not be formed are also counted here
    // second pass: kNatBigThreads polygon pairs of kNatBig lines, entry k of slot g at
    // [k * kNatBigThreads + g]
    int *btag;
    double *bvx, *bvy;
};

__device__ __forceinline__ void nat_error(const NatArgs &a)
{
    atomicAdd(&a.ctl[4], 1);
    if (a.err) atomicAdd(a.err, 1);
}

__device__ __forceinline__ double cross3(double2 a, double2 b, double2 c)
{
    return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
}

// One wave per row: first and last good node (x ascending) and the good count.
__global__ __launch_bounds__(256) void k_nat_rows(NatArgs a)
{
    const int r = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= a.R) return;
    const int k0 = a.rstart[r], k1 = a.rstart[r + 1];
    int first = 0x7fffffff, last = -1, cnt = 0;
    for (int k = k0 + lane; k < k1; k += 64)
        if (!a.bad[a.rnode[k]]) {
            first = min(first, k);
            last = max(last, k);
            ++cnt;
        }
    for (int m = 32; m >= 1; m >>= 1) {
        first = min(first, __shfl_xor(first, m, kWave));
        last = max(last, __shfl_xor(last, m, kWave));
        cnt += __shfl_xor(cnt, m, kWave);
    }
    if (lane == 0) {
        a.rows[3 * r] = cnt ? a.rnode[first] : -1;
        a.rows[3 * r + 1] = cnt ? a.rnode[last] : -1;
        a.rows[3 * r + 2] = cnt;
    }
}

// Strict hull (collinear points dropped) of the good nodes, from the row extremes (a
// node between the first and last good node of its row is not a strict vertex).  The
// extremes in the eight directions of 45-degree steps form a convex polygon of good
// nodes; candidates inside it or on its boundary are not strict vertices either, so only
// its vertices and the candidates strictly outside it enter the monotone chain, which
// runs in the axis-swapped frame (the candidates are sorted by (y, x); cross products
// change sign), is reversed to counter-clockwise and rotated to start at the lowest-x,
// lowest-y vertex.
__global__ __launch_bounds__(256) void k_nat_hull(NatArgs a)
{
    __shared__ int red[4];
    __shared__ double sv[8][4];
    __shared__ int si[8][4];
    __shared__ int oct[9], noct, cvx;
    __shared__ int pre[257];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int c = 0;
    for (int r = tid; r < a.R; r += blockDim.x) c += a.rows[3 * r + 2];
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, kWave);
    if (lane == 0) red[w] = c;
    const int nc = 2 * a.R;  // candidate slot t: row t / 2, first (even) or last (odd)
    auto cand = [&](int t) -> int {
        const int r = t >> 1;
        const int f = a.rows[3 * r], l = a.rows[3 * r + 1];
        return (t & 1) ? (l != f ? l : -1) : f;
    };
    // extremes in the directions (1,0) (1,1) (0,1) (-1,1) (-1,0) (-1,-1) (0,-1) (1,-1)
    const double dx[8] = {1, 1, 0, -1, -1, -1, 0, 1}, dy[8] = {0, 1, 1, 1, 0, -1, -1, -1};
    double bv[8];
    int bi[8];
    for (int k = 0; k < 8; ++k) { bv[k] = -INFINITY; bi[k] = 0x7fffffff; }
    for (int t = tid; t < nc; t += blockDim.x) {
        const int p = cand(t);
        if (p < 0) continue;
        const double2 P = a.xy[p];
        for (int k = 0; k < 8; ++k) {
            const double v = dx[k] * P.x + dy[k] * P.y;
            if (v > bv[k] || (v == bv[k] && p < bi[k])) { bv[k] = v; bi[k] = p; }
        }
    }
    for (int k = 0; k < 8; ++k) {
        for (int m = 32; m >= 1; m >>= 1) {
            const double v2 = __shfl_xor(bv[k], m, kWave);
            const int i2 = __shfl_xor(bi[k], m, kWave);
            if (v2 > bv[k] || (v2 == bv[k] && i2 < bi[k])) { bv[k] = v2; bi[k] = i2; }
        }
        if (lane == 0) { sv[k][w] = bv[k]; si[k][w] = bi[k]; }
    }
    __syncthreads();
    if (tid == 0) {
        int n = 0;
        for (int k = 0; k < 8; ++k) {
            double v = sv[k][0];
            int i = si[k][0];
            for (int q = 1; q < 4; ++q)
                if (sv[k][q] > v || (sv[k][q] == v && si[k][q] < i)) { v = sv[k][q]; i = si[k][q]; }
            if (i == 0x7fffffff) continue;
            if (n == 0 || oct[n - 1] != i) oct[n++] = i;
        }
        while (n > 1 && oct[n - 1] == oct[0]) --n;
        int convex = n >= 3;
        for (int e = 0; e < n && convex; ++e)
            if (!(cross3(a.xy[oct[e]], a.xy[oct[(e + 1) % n]], a.xy[oct[(e + 2) % n]]) > 0.0))
                convex = 0;
        noct = n;
        cvx = convex;
    }
    __syncthreads();
    // survivors: the octagon's vertices and candidates strictly outside it (all, when it
    // is not a proper convex polygon); counted per thread over a contiguous chunk
    const int per = (nc + 255) / 256, t0 = tid * per, t1 = min(nc, t0 + per);
    auto survives = [&](int p) -> bool {
        if (!cvx) return true;
        for (int e = 0; e < noct; ++e)
            if (oct[e] == p) return true;
        const double2 P = a.xy[p];
        for (int e = 0; e < noct; ++e)
            if (cross3(a.xy[oct[e]], a.xy[oct[e + 1 < noct ? e + 1 : 0]], P) < 0.0) return true;
        return false;
    };
    int mine = 0;
    for (int t = t0; t < t1; ++t) {
        const int p = cand(t);
        if (p >= 0 && survives(p)) ++mine;
    }
    pre[tid + 1] = mine;
    __syncthreads();
    if (tid == 0) {
        pre[0] = 0;
        for (int k = 1; k <= 256; ++k) pre[k] += pre[k - 1];
    }
    __syncthreads();
    int *P = a.chain;
    for (int t = t0, o = pre[tid]; t < t1; ++t) {
        const int p = cand(t);
        if (p >= 0 && survives(p)) P[o++] = p;
    }
    __syncthreads();
    if (tid != 0) return;
    const int ng = red[0] + red[1] + red[2] + red[3];
    const int np = pre[256];
    int *H = a.chain + np;       // chain stack (at most 2 np entries)
    int k = 0;
    for (int t = 0; t < np; ++t) {
        while (k >= 2 && cross3(a.xy[H[k - 2]], a.xy[H[k - 1]], a.xy[P[t]]) >= 0.0) --k;
        H[k++] = P[t];
    }
    for (int t = np - 2, lo = k + 1; t >= 0; --t) {
        while (k >= lo && cross3(a.xy[H[k - 2]], a.xy[H[k - 1]], a.xy[P[t]]) >= 0.0) --k;
        H[k++] = P[t];
    }
    int h = np >= 3 ? k - 1 : 0;
    if (h < 3) h = 0;
    int s0 = 0;
    for (int t = 1; t < h; ++t) {
        const double2 p = a.xy[H[t]], b = a.xy[H[s0]];
        if (p.x < b.x || (p.x == b.x && p.y < b.y)) s0 = t;
    }
    for (int t = 0; t < h; ++t) a.hull[t] = H[((s0 - t) % h + h) % h];
    a.ctl[0] = h;
    a.ctl[1] = ng;
    a.ctl[2] = 0;
    a.ctl[3] = 0;
    a.ctl[4] = 0;
    a.ctl[5] = 0;
    a.ctl[6] = 0;
}

// Inside = strictly left of every hull edge (relative tolerance 1e-12).  The hull is
// convex and counter-clockwise from v0 = hull[0]: after the two edges at v0, a binary
// search over the fan from v0 finds the wedge (v0, v_k, v_k+1) holding q, and only the
// edges around it (k - 1, k, k + 1) can be the one q is not strictly left of.
__global__ __launch_bounds__(256) void k_nat_classify(NatArgs a)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool mine = q < a.nn && a.bad[q];
    const int h = a.ctl[0];
    bool inside = mine && h >= 3 && a.ctl[1] >= 3;
    const double2 Q = mine ? a.xy[q] : make_double2(0.0, 0.0);
    auto left = [&](int e) {
        const double2 A = a.xy[a.hull[e]], C = a.xy[a.hull[e + 1 < h ? e + 1 : 0]];
        const double ex = C.x - A.x, ey = C.y - A.y, px = Q.x - A.x, py = Q.y - A.y;
        const double cr = ex * py - ey * px;
        return cr > 1e-12 * sqrt(ex * ex + ey * ey) * sqrt(px * px + py * py);
    };
    if (inside) inside = left(0) && left(h - 1);
    if (inside) {
        const double2 V0 = a.xy[a.hull[0]];
        int lo = 1, hi = h - 1;  // cross(v0, v_lo, q) > 0 >= cross(v0, v_hi, q)
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (cross3(V0, a.xy[a.hull[mid]], Q) > 0.0) lo = mid;
            else hi = mid;
        }
        inside = left(lo) && left(lo - 1) && left(lo + 1);
    }
    // one atomic per wave and list (every lane reaches the ballots)
    const unsigned long long m_in = __ballot(mine && inside), m_out = __ballot(mine && !inside);
    const int lane = threadIdx.x & 63;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int b_in = 0, b_out = 0;
    if (lane == 0) {
        if (m_in) b_in = atomicAdd(&a.ctl[2], __popcll(m_in));
        if (m_out) b_out = atomicAdd(&a.ctl[3], __popcll(m_out));
    }
    b_in = __shfl(b_in, 0, kWave);
    b_out = __shfl(b_out, 0, kWave);
    if (!mine) return;
    if (inside) a.lin[b_in + __popcll(m_in & below)] = q;
    else a.lhull[b_out + __popcll(m_out & below)] = q;
}

// ---- per-thread polygons in LDS ------------------------------------------------------
template <int ST>
struct Poly {        // entry k at [k * ST]: a column of [lines][threads]
    int *tag;
    double *vx, *vy;
};

struct Line {
    double nx, ny, c;
};

// Site lookups: coordinates of the site a tag names, relative to q.  SitesG reads the
// node table (tags are node indices), SitesL an LDS copy of one cell's sites (tags are
// local indices); both subtract q the same way.
struct SitesG {
    const double2 *xy;
    double2 Q;
    __device__ double2 rel(int t) const
    {
        const double2 P = xy[t];
        return make_double2(P.x - Q.x, P.y - Q.y);
    }
};
struct SitesL {
    const double *sx, *sy;
    __device__ double2 rel(int t) const { return make_double2(sx[t], sy[t]); }
};

// Line of a tag, coordinates relative to q; pi = p_i - q for region tags.
template <class S>
__device__ __forceinline__ Line nat_line_s(const S &st, double B, int tag, double2 pi)
{
    Line L;
    if (tag < 0) {
        const int k = -1 - tag;
        L.nx = k == 0 ? 1.0 : (k == 2 ? -1.0 : 0.0);
        L.ny = k == 1 ? 1.0 : (k == 3 ? -1.0 : 0.0);
        L.c = B;
    } else if (tag & kRegion) {
        const double2 P = st.rel(tag & ~kRegion);
        const double px = P.x, py = P.y;
        L.nx = px - pi.x;
        L.ny = py - pi.y;
        L.c = 0.5 * ((px * px + py * py) - (pi.x * pi.x + pi.y * pi.y));
    } else {
        const double2 P = st.rel(tag);
        const double px = P.x, py = P.y;
        L.nx = px;
        L.ny = py;
        L.c = 0.5 * (px * px + py * py);
    }
    return L;
}

__device__ __forceinline__ Line nat_line(const NatArgs &a, double2 Q, double B, int tag,
                                         double2 pi)
{
    return nat_line_s(SitesG{a.xy, Q}, B, tag, pi);
}

__device__ __forceinline__ void nat_meet(const Line &l1, const Line &l2, double &x, double &y)
{
    const double det = l1.nx * l2.ny - l1.ny * l2.nx;
    x = (l1.c * l2.ny - l2.c * l1.ny) / det;
    y = (l1.nx * l2.c - l2.nx * l1.c) / det;
}

#define PT(P, k) (P).tag[(k) * ST]
#define PX(P, k) (P).vx[(k) * ST]
#define PY(P, k) (P).vy[(k) * ST]

// Clip P (m entries, room for MAXM) with line M (tag mt): keep M.n . x <= M.c.  Returns
// the new size, 0 when nothing is left, -1 when the vertices outside M do not form one
// run (the polygon lost its convexity numerically), -2 when MAXM entries are too few.
template <int ST, int MAXM, class S>
__device__ int nat_clip_s(const S &st, double B, double2 pi, Poly<ST> P, int m, const Line &M, int mt)
{
    auto outside = [&](int k) {
        const double u = M.nx * PX(P, k), v = M.ny * PY(P, k);
        const double d = u + v - M.c;
        return d > 1e-12 * (fabs(u) + fabs(v) + fabs(M.c));
    };
    int len, aa;
    if constexpr (MAXM <= 32) {  // the outside set as a bit mask
        unsigned mask = 0u;
        for (int k = 0; k < m; ++k)
            if (outside(k)) mask |= 1u << k;
        if (!mask) return m;
        len = __popc(mask);
        if (len == m) return 0;
        const unsigned full = m == 32 ? 0xffffffffu : ((1u << m) - 1u);
        const unsigned prev = ((mask << 1) | (mask >> (m - 1))) & full;  // bit k: out[k - 1]
        aa = __ffs(mask & ~prev) - 1;                                    // first of the run
        int run = 0;
        while ((mask >> ((aa + run) % m)) & 1u) ++run;
        if (run != len) return -1;
    } else {
        len = 0;
        aa = -1;
        bool prev = outside(m - 1);
        for (int k = 0; k < m; ++k) {
            const bool o = outside(k);
            if (o) {
                ++len;
                if (!prev && aa < 0) aa = k;  // first vertex of the outside run
            }
            prev = o;
        }
        if (len == 0) return m;
        if (len == m) return 0;
        int run = 0;
        while (run < m && outside((aa + run) % m)) ++run;
        if (run != len) return -1;
    }
    // delete entries aa+1 .. aa+len-1 (cyclic), insert M after aa
    int ins;  // slot of M
    int nm;
    if (len >= 2) {
        ins = aa + 1 < m ? aa + 1 : 0;
        PT(P, ins) = mt;
        const int d = len - 2;
        nm = m - d;
        if (d > 0) {
            const int s = ins + 1 < m ? ins + 1 : 0;
            if (s + d <= m) {
                for (int k = s; k + d < m; ++k) {
                    PT(P, k) = PT(P, k + d); PX(P, k) = PX(P, k + d); PY(P, k) = PY(P, k + d);
                }
            } else {
                const int sh = s + d - m;  // entries [sh, s) are kept, moved to 0
                for (int k = 0; k < s - sh; ++k) {
                    PT(P, k) = PT(P, k + sh); PX(P, k) = PX(P, k + sh); PY(P, k) = PY(P, k + sh);
                }
            }
        }
    } else {
        if (m + 1 > MAXM) return -2;
        for (int k = m; k > aa + 1; --k) {
            PT(P, k) = PT(P, k - 1); PX(P, k) = PX(P, k - 1); PY(P, k) = PY(P, k - 1);
        }
        ins = aa + 1;
        PT(P, ins) = mt;
        nm = m + 1;
    }
    // locate M again (the deletion may have moved it) and refresh the two new vertices
    int km = 0;
    while (km < nm && PT(P, km) != mt) ++km;
    if (km == nm) return -1;
    const int kp = km > 0 ? km - 1 : nm - 1, kn = km + 1 < nm ? km + 1 : 0;
    const Line Lp = nat_line_s(st, B, PT(P, kp), pi), Ln = nat_line_s(st, B, PT(P, kn), pi);
    double x, y;
    nat_meet(Lp, M, x, y);
    PX(P, kp) = x; PY(P, kp) = y;
    nat_meet(M, Ln, x, y);
    PX(P, km) = x; PY(P, km) = y;
    return nm;
}

template <int ST, int MAXM>
__device__ __forceinline__ int nat_clip(const NatArgs &a, double2 Q, double B, double2 pi, Poly<ST> P,
                                        int m, const Line &M, int mt)
{
    return nat_clip_s<ST, MAXM>(SitesG{a.xy, Q}, B, pi, P, m, M, mt);
}

template <int ST>
__device__ double nat_radius(Poly<ST> P, int m)
{
    double r2 = 0.0;
    for (int k = 0; k < m; ++k) r2 = fmax(r2, PX(P, k) * PX(P, k) + PY(P, k) * PY(P, k));
    return sqrt(r2);
}

template <int ST>
__device__ double nat_area(Poly<ST> P, int m)
{
    double s = 0.0;
    for (int k = 0; k < m; ++k) {
        const int k0 = k > 0 ? k - 1 : m - 1;
        s += PX(P, k0) * PY(P, k) - PX(P, k) * PY(P, k0);
    }
    return 0.5 * s;
}

// Sibson weights of one inside node q with polygons V (the cell of q) and R (the part
// of it taken from one natural neighbour), MAXM lines each.  Returns 0 (values written),
// -1 (the cell could not be formed) or -2 (MAXM lines are too few; nothing written).
template <int ST, int MAXM, int RING_CAP = 1 << 30>
__device__ int nat_sibson_node(const NatArgs &a, int q, Poly<ST> V, Poly<ST> R)
{
    const double2 Q = a.xy[q];
    const double B = 1e6 * (a.h * (a.BX + a.BY) + 1.0);
    const double2 zero = make_double2(0.0, 0.0);
    // initial square: lines x <= B, y <= B, -x <= B, -y <= B (counter-clockwise)
    int m = 4;
    for (int k = 0; k < 4; ++k) PT(V, k) = -1 - k;
    for (int k = 0; k < 4; ++k) {
        double x, y;
        nat_meet(nat_line(a, Q, B, -1 - k, zero), nat_line(a, Q, B, -1 - ((k + 1) & 3), zero),
                 x, y);
        PX(V, k) = x; PY(V, k) = y;
    }
    double rho = nat_radius(V, m);
    int bx = (int)floor((Q.x - a.ox) / a.h), by = (int)floor((Q.y - a.oy) / a.h);
    bx = bx < 0 ? 0 : (bx >= a.BX ? a.BX - 1 : bx);
    by = by < 0 ? 0 : (by >= a.BY ? a.BY - 1 : by);
    const int rmax = max(max(bx, a.BX - 1 - bx), max(by, a.BY - 1 - by));
    int fail = 0;
    // the good nodes of bucket (i, j) cut V(q)
    auto bucket = [&](int i, int j) {
        if (fail || i < 0 || i >= a.BX || j < 0 || j >= a.BY) return;
        const int b = j * a.BX + i;
        for (int s = a.bstart[b]; s < a.bstart[b + 1]; ++s) {
            const int p = a.bnode[s];
            if (p == q || a.bad[p]) continue;
            const double2 P = a.xy[p];
            const double px = P.x - Q.x, py = P.y - Q.y;
            if (px * px + py * py >= 4.0 * rho * rho) continue;  // cannot reach the cell
            const Line M{px, py, 0.5 * (px * px + py * py)};
            m = nat_clip<ST, MAXM>(a, Q, B, zero, V, m, M, p);
            if (m <= 0) { fail = m == -2 ? -2 : -1; return; }
            rho = nat_radius(V, m);
        }
    };
    for (int r = 0; r <= rmax && !fail; ++r) {
        // rings 0 .. r-1 done: every node left is more than (r - 1) h from q
        if (r > 0 && (double)(r - 1) * a.h >= 2.0 * rho) break;
        if (r > RING_CAP) {  // a cell still open this far out reaches across a bad cluster
            for (int k = 0; k < m; ++k)
                if (PT(V, k) < 0) return -2;
        }
        if (r == 0) {
            bucket(bx, by);
            continue;
        }
        for (int i = bx - r; i <= bx + r; ++i) { bucket(i, by - r); bucket(i, by + r); }
        for (int j = by - r + 1; j <= by + r - 1; ++j) { bucket(bx - r, j); bucket(bx + r, j); }
    }
    if (fail == -2) return -2;
    for (int k = 0; k < m && !fail; ++k)
        if (PT(V, k) < 0) fail = -1;
    double2 u = make_double2(NAN, NAN);
    double4 f = make_double4(NAN, NAN, NAN, NAN);
    if (!fail) {
        double tot = 0.0;
        double2 su = make_double2(0.0, 0.0);
        double4 sf = make_double4(0.0, 0.0, 0.0, 0.0);
        for (int k = 0; k < m && !fail; ++k) {
            const int ti = PT(V, k);
            const double2 Pi = a.xy[ti];
            const double2 pi = make_double2(Pi.x - Q.x, Pi.y - Q.y);
            for (int l = 0; l < m; ++l) {
                PT(R, l) = PT(V, l); PX(R, l) = PX(V, l); PY(R, l) = PY(V, l);
            }
            // the other neighbours' bisectors outward from k along the cell (k + 1, k - 1,
            // k + 2, ...): the adjacent ones cut the region down at once, so the rest test
            // a few vertices each (the intersection does not depend on the order)
            int mr = m;
            for (int d = 1; d < m && mr > 0; ++d) {
                const int l = (k + ((d & 1) ? (d + 1) / 2 : m - d / 2)) % m;
                const int tj = PT(V, l);
                const double2 Pj = a.xy[tj];
                const double pjx = Pj.x - Q.x, pjy = Pj.y - Q.y;
                const Line M{pjx - pi.x, pjy - pi.y,
                             0.5 * ((pjx * pjx + pjy * pjy) - (pi.x * pi.x + pi.y * pi.y))};
                mr = nat_clip<ST, MAXM>(a, Q, B, pi, R, mr, M, tj | kRegion);
                if (mr < 0) { fail = mr; break; }
            }
            if (fail) break;
            const double w = mr > 0 ? nat_area(R, mr) : 0.0;
            tot += w;
            const double2 uv = a.U[ti];
            su.x += w * uv.x; su.y += w * uv.y;
            if (a.F) {
                const double4 fv = a.F[ti];
                sf.x += w * fv.x; sf.y += w * fv.y; sf.z += w * fv.z; sf.w += w * fv.w;
            }
        }
        if (fail == -2) return -2;
        if (!fail) {
            const double it = 1.0 / tot;
            u = make_double2(su.x * it, su.y * it);
            f = make_double4(sf.x * it, sf.y * it, sf.z * it, sf.w * it);
        }
    }
    a.U[q] = u;
    if (a.F) a.F[q] = f;
    return fail;
}

// First pass: one thread per inside node, polygons of kNatMax lines in LDS; a node whose
// cell needs more lines goes to the second pass's list.
__global__ __launch_bounds__(kNatThreads) void k_nat_sibson(NatArgs a)
{
    __shared__ int s_tag[2][kNatMax * kNatThreads];
    __shared__ double s_vx[2][kNatMax * kNatThreads], s_vy[2][kNatMax * kNatThreads];
    const int t = threadIdx.x;
    const Poly<kNatThreads> V{&s_tag[0][t], &s_vx[0][t], &s_vy[0][t]};
    const Poly<kNatThreads> R{&s_tag[1][t], &s_vx[1][t], &s_vy[1][t]};
    const int n_in = a.ctl[2];
    if (n_in <= kNatAllWave) return;   // few nodes: one wave each (lower latency per cell)
    for (int li = blockIdx.x * kNatThreads + t; li < n_in; li += gridDim.x * kNatThreads) {
        const int q = a.lin[li];
        const int r = nat_sibson_node<kNatThreads, kNatMax, kNatRingCap>(a, q, V, R);
        if (r == -2) a.lbig[atomicAdd(&a.ctl[5], 1)] = q;
        else if (r < 0) nat_error(a);
    }
}

// ---- second pass: one wave per node whose cell has many natural neighbours ------------
// The cell V(q) as two LDS buffers of cyclic line lists; a clip is done by the wave: every
// lane tests its vertices against the line, the single run of outside vertices [aa, aa +
// len) is located by ballots, and the survivors are copied to the other buffer rotated
// so that the list reads (line aa, M, line aa + len, ..., line aa - 1); the two new
// vertices are the meets of M with its neighbours.  Same tests and vertices as nat_clip.
struct NatWaveLds {
    int tag[2][kNatWaveV];
    double vx[2][kNatWaveV], vy[2][kNatWaveV];
    double sx[kNatWaveV], sy[kNatWaveV];         // the cell's sites relative to q
    union {   // the candidates while the cell is built, then the lanes' region polygons
        struct {
            int cand[kNatWaveCand];
            double cx[kNatWaveCand], cy[kNatWaveCand];   // candidates relative to q
        };
        struct {
            int rtag[kNatWaveR * 64];
            double rvx[kNatWaveR * 64], rvy[kNatWaveR * 64];
        };
    };
};
static_assert(kNatWaveCand >= kNatWaveR * 64, "region polygons share the candidates' LDS");
static_assert(sizeof(NatWaveLds) <= 40 * 1024, "four wave-pass workgroups per CU");

__device__ __forceinline__ bool nat_outside(double x, double y, const Line &M)
{
    const double u = M.nx * x, v = M.ny * y;
    const double d = u + v - M.c;
    return d > 1e-12 * (fabs(u) + fabs(v) + fabs(M.c));
}

__device__ __forceinline__ double wave_max_d(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, kWave));
    return v;
}

__device__ __forceinline__ int wave_sum_i(int v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Returns the new size (m when M does not cut, 0 when nothing is left), -1 (the outside
// vertices are not one run) or -2 (more than kNatWaveV lines); cur flips on a cut.
__device__ int nat_wclip(const NatArgs &a, double2 Q, double B, NatWaveLds &L, int &cur, int m,
                         const Line &M, int mt, int lane)
{
    int len = 0, nstart = 0, aa = -1;
    for (int base = 0; base < m; base += kWave) {
        const int k = base + lane;
        bool o = false, op = false;
        if (k < m) {
            const int kp = k > 0 ? k - 1 : m - 1;
            o = nat_outside(L.vx[cur][k], L.vy[cur][k], M);
            op = nat_outside(L.vx[cur][kp], L.vy[cur][kp], M);
        }
        const unsigned long long bo = __ballot(o), bs = __ballot(o && !op);
        len += __popcll(bo);
        nstart += __popcll(bs);
        if (bs && aa < 0) aa = base + __ffsll((long long)bs) - 1;
    }
    if (len == 0) return m;
    if (len == m) return 0;
    if (nstart != 1) return -1;
    const int nm = m - len + 2;
    if (nm > kNatWaveV) return -2;
    const int nx = 1 - cur;
    for (int j = lane; j < m - len; j += kWave) {
        const int s = (aa + len + j) % m;
        L.tag[nx][2 + j] = L.tag[cur][s];
        L.vx[nx][2 + j] = L.vx[cur][s];
        L.vy[nx][2 + j] = L.vy[cur][s];
    }
    if (lane == 0) {
        const double2 zero = make_double2(0.0, 0.0);
        const int ta = L.tag[cur][aa], tn = L.tag[cur][(aa + len) % m];
        double x, y;
        L.tag[nx][0] = ta;
        L.tag[nx][1] = mt;
        nat_meet(nat_line(a, Q, B, ta, zero), M, x, y);
        L.vx[nx][0] = x; L.vy[nx][0] = y;
        nat_meet(M, nat_line(a, Q, B, tn, zero), x, y);
        L.vx[nx][1] = x; L.vy[nx][1] = y;
    }
    wave_sync_lds();
    cur = nx;
    return nm;
}

// Bucket t (< 8 r, or the one bucket of ring 0) of ring r around (bx, by), in the order
// nat_sibson_node visits them.
// ring of bucket T of the ring sequence (ring r >= 1 holds T = 1 + 4 r (r - 1) .. 4 r (r + 1))
__device__ __forceinline__ int bucket_ring(long T)
{
    if (T <= 0) return 0;
    int r = (int)((1.0 + sqrt((double)T)) * 0.5);
    while (1 + 4L * r * (r + 1) <= T) ++r;
    while (r > 1 && 1 + 4L * r * (r - 1) > T) --r;
    return r;
}

__device__ __forceinline__ void ring_bucket(int r, int t, int bx, int by, int &i, int &j)
{
    if (r == 0) { i = bx; j = by; return; }
    const int n1 = 2 * (2 * r + 1);
    if (t < n1) { i = bx - r + t / 2; j = (t & 1) ? by + r : by - r; return; }
    t -= n1;
    j = by - r + 1 + t / 2;
    i = (t & 1) ? bx + r : bx - r;
}

// Sibson weights of inside node q by one wave: the cell nat_sibson_node builds (the same
// bisectors; sites that cut nothing are dropped before the wave-parallel clips, and a
// vertex is always the meet of its two lines, so the order does not change it), then lane k the region of natural
// neighbour k: the square clipped by q's bisector with p_k and by p_k's bisectors with
// p_(k+1) and p_(k-1) first, then by the others outward along the cell, so it stays a few
// lines.  Returns 0, -1 (the cell could not be formed) or -2 (beyond the LDS limits).
__device__ int nat_sibson_wave(const NatArgs &a, int q, NatWaveLds &L, int lane)
{
    const double2 Q = a.xy[q];
    const double B = 1e6 * (a.h * (a.BX + a.BY) + 1.0);
    const double2 zero = make_double2(0.0, 0.0);
    int cur = 0, m = 4;
    if (lane < 4) {
        double x, y;
        nat_meet(nat_line(a, Q, B, -1 - lane, zero), nat_line(a, Q, B, -1 - ((lane + 1) & 3), zero), x, y);
        L.tag[0][lane] = -1 - lane;
        L.vx[0][lane] = x;
        L.vy[0][lane] = y;
    }
    wave_sync_lds();
    // the cell's radius and its supports h(d) = max_v d . v (relative to q) in the eight
    // directions +-x, +-y, (+-1, +-1): a site p cuts the cell only if h(p) > |p|^2 / 2, and
    // h is sublinear, so with p = alpha a + beta b over the two of those directions around
    // it (alpha, beta >= 0), h(p) <= alpha h(a) + beta h(b) -- an octagon around the cell,
    // much tighter than its box for the long thin cells of nodes next to a hole
    double rho, hx[2], hy[2], hd[4];   // hx[sign x], hy[sign y], hd[2 (y < 0) + (x < 0)]
    auto shape = [&]() {
        double r2 = 0.0, e[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) e[i] = -INFINITY;
        for (int k = lane; k < m; k += kWave) {
            const double x = L.vx[cur][k], y = L.vy[cur][k];
            r2 = fmax(r2, x * x + y * y);
            e[0] = fmax(e[0], x); e[1] = fmax(e[1], -x); e[2] = fmax(e[2], y); e[3] = fmax(e[3], -y);
            e[4] = fmax(e[4], x + y); e[5] = fmax(e[5], y - x); e[6] = fmax(e[6], x - y);
            e[7] = fmax(e[7], -x - y);
        }
        rho = sqrt(wave_max_d(r2));
        hx[0] = wave_max_d(e[0]); hx[1] = wave_max_d(e[1]);
        hy[0] = wave_max_d(e[2]); hy[1] = wave_max_d(e[3]);
        hd[0] = wave_max_d(e[4]); hd[1] = wave_max_d(e[5]); hd[2] = wave_max_d(e[6]);
        hd[3] = wave_max_d(e[7]);
    };
    auto reaches = [&](double px, double py) {
        const double d2 = px * px + py * py;
        if (!(d2 < 4.0 * rho * rho)) return false;
        const int sx = px < 0.0, sy = py < 0.0;
        const double ax = fabs(px), ay = fabs(py), sd = hd[2 * sy + sx];
        const double h = ax >= ay ? (ax - ay) * hx[sx] + ay * sd : (ay - ax) * hy[sy] + ax * sd;
        return h > 0.5 * d2;
    };
    auto cuts = reaches;
    shape();
    int bx = (int)floor((Q.x - a.ox) / a.h), by = (int)floor((Q.y - a.oy) / a.h);
    bx = bx < 0 ? 0 : (bx >= a.BX ? a.BX - 1 : bx);
    by = by < 0 ? 0 : (by >= a.BY ? a.BY - 1 : by);
    const int rmax = max(max(bx, a.BX - 1 - bx), max(by, a.BY - 1 - by));
    // the ring buckets as one sequence (ring 0 is bucket 0, ring r >= 1 buckets
    // 1 + 4 r (r - 1) ..), in chunks of 64 buckets, then of kNatLaneB x 64 (lane l takes
    // buckets l, l + 64, ...): fewer round trips for the cells that reach far rings.  A
    // chunk may run past the ring where the scan could stop; the extra sites cut nothing
    // and the filter drops them.
    const long nbt = 1 + 4L * rmax * (rmax + 1);
    for (long T0 = 0, cs = kWave; T0 < nbt; T0 += cs, cs = kNatLaneB * kWave) {
        const int r0 = bucket_ring(T0);
        if (r0 > 0 && (double)(r0 - 1) * a.h >= 2.0 * rho) break;
        const int nb_l = cs / kWave;
        int s0[kNatLaneB], s1[kNatLaneB];
#pragma unroll
        for (int j = 0; j < kNatLaneB; ++j) {
            s0[j] = s1[j] = 0;
            const long T = T0 + lane + (long)kWave * j;
            if (j < nb_l && T < nbt) {
                const int r = bucket_ring(T);
                int i, jj;
                ring_bucket(r, (int)(r == 0 ? 0 : T - (1 + 4L * r * (r - 1))), bx, by, i, jj);
                if (i >= 0 && i < a.BX && jj >= 0 && jj < a.BY) {
                    s0[j] = a.bstart[jj * a.BX + i];
                    s1[j] = a.bstart[jj * a.BX + i + 1];
                }
            }
        }
        // the first kNatPre nodes of each bucket with every load in flight, in registers;
        // nodes past them (dense buckets) by a loop
        int pn[kNatLaneB][kNatPre];
        double pxs[kNatLaneB][kNatPre], pys[kNatLaneB][kNatPre];
#pragma unroll
        for (int j = 0; j < kNatLaneB; ++j)
#pragma unroll
            for (int u = 0; u < kNatPre; ++u) pn[j][u] = s0[j] + u < s1[j] ? a.bnode[s0[j] + u] : -1;
        int c = 0;
#pragma unroll
        for (int j = 0; j < kNatLaneB; ++j)
#pragma unroll
            for (int u = 0; u < kNatPre; ++u) {
                const int p = pn[j][u] < 0 ? q : pn[j][u];
                const double2 P = a.xy[p];
                const int bd = a.bad[p];
                pxs[j][u] = P.x - Q.x;
                pys[j][u] = P.y - Q.y;
                if (p == q || bd || !cuts(pxs[j][u], pys[j][u])) pn[j][u] = -1;
                c += pn[j][u] >= 0;
            }
        auto rest = [&](auto &&take) {
            for (int j = 0; j < kNatLaneB; ++j)
                for (int sx = s0[j] + kNatPre; sx < s1[j]; ++sx) {
                    const int p = a.bnode[sx];
                    if (p == q || a.bad[p]) continue;
                    const double px = a.xy[p].x - Q.x, py = a.xy[p].y - Q.y;
                    if (cuts(px, py)) take(p, px, py);
                }
        };
        rest([&](int, double, double) { ++c; });
        int off = c;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {   // inclusive scan over lanes
            const int o = __shfl_up(off, d, kWave);
            if (lane >= d) off += o;
        }
        const int total = __shfl(off, kWave - 1, kWave);
        if (total > kNatWaveCand) return -2;
        off -= c;
        auto put = [&](int p, double px, double py) {
            L.cand[off] = p;
            L.cx[off] = px;
            L.cy[off] = py;
            ++off;
        };
#pragma unroll
        for (int j = 0; j < kNatLaneB; ++j)
#pragma unroll
            for (int u = 0; u < kNatPre; ++u)
                if (pn[j][u] >= 0) put(pn[j][u], pxs[j][u], pys[j][u]);
        rest(put);
        wave_sync_lds();
        for (int ci = 0; ci < total; ++ci) {
            const int p = L.cand[ci];
            const double px = L.cx[ci], py = L.cy[ci];
            if (px * px + py * py >= 4.0 * rho * rho) continue;   // the cell shrank since
            if (!reaches(px, py)) continue;
            const Line M{px, py, 0.5 * (px * px + py * py)};
            const int m0 = m;
            m = nat_wclip(a, Q, B, L, cur, m, M, p, lane);
            if (m <= 0) return m == -2 ? -2 : -1;
            if (m != m0 || L.tag[cur][1] == p) shape();
        }
        wave_sync_lds();
    }
    // a cell still touching the initial square could not be formed
    bool open = false;
    for (int k = lane; k < m; k += kWave) open = open || L.tag[cur][k] < 0;
    if (__ballot(open)) return -1;
    // the cell's sites in LDS (local tags 0 .. m-1)
    for (int k = lane; k < m; k += kWave) {
        const double2 P = a.xy[L.tag[cur][k]];
        L.sx[k] = P.x - Q.x;
        L.sy[k] = P.y - Q.y;
    }
    wave_sync_lds();
    const SitesL sl{L.sx, L.sy};
    // regions, lane k for natural neighbours k, k + 64, ...
    const Poly<kWave> R{&L.rtag[lane], &L.rvx[lane], &L.rvy[lane]};
    double tot = 0.0;
    double2 su = make_double2(0.0, 0.0);
    double4 sf = make_double4(0.0, 0.0, 0.0, 0.0);
    int fail = 0;
    for (int k = lane; k < m && !fail; k += kWave) {
        const int tk = L.tag[cur][k];
        const double2 pi = make_double2(L.sx[k], L.sy[k]);
        int mr = 4;
        for (int e = 0; e < 4; ++e) {
            double x, y;
            nat_meet(nat_line_s(sl, B, -1 - e, pi), nat_line_s(sl, B, -1 - ((e + 1) & 3), pi), x, y);
            R.tag[e * kWave] = -1 - e;
            R.vx[e * kWave] = x;
            R.vy[e * kWave] = y;
        }
        // q's bisector with p_k, then outward: p_k's bisectors with p_(k+d), p_(k-d) and
        // q's with them (local tags)
        mr = nat_clip_s<kWave, kNatWaveR>(sl, B, pi, R, mr, nat_line_s(sl, B, k, pi), k);
        for (int d = 1; d < m && mr > 0; ++d) {
            const int l = (k + ((d & 1) ? (d + 1) / 2 : m - d / 2)) % m;
            mr = nat_clip_s<kWave, kNatWaveR>(sl, B, pi, R, mr, nat_line_s(sl, B, l | kRegion, pi), l | kRegion);
            if (mr > 0) mr = nat_clip_s<kWave, kNatWaveR>(sl, B, pi, R, mr, nat_line_s(sl, B, l, pi), l);
        }
        if (mr < 0) { fail = mr; break; }
        const double w = mr > 0 ? nat_area(R, mr) : 0.0;
        tot += w;
        const double2 uv = a.U[tk];
        su.x += w * uv.x; su.y += w * uv.y;
        if (a.F) {
            const double4 fv = a.F[tk];
            sf.x += w * fv.x; sf.y += w * fv.y; sf.z += w * fv.z; sf.w += w * fv.w;
        }
    }
    const unsigned long long bf = __ballot(fail != 0), b2 = __ballot(fail == -2);
    if (b2) return -2;
    if (bf) return -1;
    tot = wave_sum(tot);
    su.x = wave_sum(su.x); su.y = wave_sum(su.y);
    if (a.F) { sf.x = wave_sum(sf.x); sf.y = wave_sum(sf.y); sf.z = wave_sum(sf.z); sf.w = wave_sum(sf.w); }
    if (lane == 0) {
        const double it = 1.0 / tot;
        a.U[q] = make_double2(su.x * it, su.y * it);
        if (a.F) a.F[q] = make_double4(sf.x * it, sf.y * it, sf.z * it, sf.w * it);
    }
    return 0;
}

// One wave per node of the big list; nodes beyond the LDS limits go to lbig2.
__global__ __launch_bounds__(64) void k_nat_sibson_wave(NatArgs a)
{
    __shared__ NatWaveLds L;
    const int lane = threadIdx.x;
    // every inside node when they are few (the thread pass did not run), else the ones
    // the thread pass left open
    const bool all = a.ctl[2] <= kNatAllWave;
    const int nb = all ? a.ctl[2] : a.ctl[5];
    const int *list = all ? a.lin : a.lbig;
    for (int li = blockIdx.x; li < nb; li += gridDim.x) {
        const int q = list[li];
        const int r = nat_sibson_wave(a, q, L, lane);
        if (lane == 0) {
            if (r == -2) a.lbig2[atomicAdd(&a.ctl[6], 1)] = q;
            else if (r != 0) nat_error(a);
        }
        __syncthreads();   // the next node reuses the LDS
    }
}

// Third pass over the nodes beyond k_nat_sibson_wave's limits: kNatBigThreads threads,
// polygons of kNatBig lines in device memory.
__global__ __launch_bounds__(64) void k_nat_sibson_big(NatArgs a)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;  // < kNatBigThreads
    const size_t half = (size_t)kNatBig * kNatBigThreads;
    const Poly<kNatBigThreads> V{a.btag + g, a.bvx + g, a.bvy + g};
    const Poly<kNatBigThreads> R{a.btag + half + g, a.bvx + half + g, a.bvy + half + g};
    const int nb = a.ctl[6];
    for (int li = g; li < nb; li += kNatBigThreads)
        if (nat_sibson_node<kNatBigThreads, kNatBig>(a, a.lbig2[li], V, R) != 0) nat_error(a);
}

#undef PT
#undef PX
#undef PY

// ---- hull-list nodes: one thread each, scans over rings of buckets -----------------
__device__ __forceinline__ bool on_seg(double2 A, double2 C, double2 P, double &s)
{
    const double ex = C.x - A.x, ey = C.y - A.y, px = P.x - A.x, py = P.y - A.y;
    const double L2 = ex * ex + ey * ey;
    const double cr = ex * py - ey * px;
    if (fabs(cr) > 1e-12 * sqrt(L2) * sqrt(px * px + py * py)) return false;
    s = (px * ex + py * ey) / L2;
    return s >= 0.0 && s <= 1.0;
}

// Visits the good nodes of the buckets in rings around point X, ring by ring; before ring
// r (r >= 1) stop(D) is asked with D = (r - 1) h: every node not yet visited is more than
// D from X.
template <class Visit, class Stop>
__device__ void ring_scan(const NatArgs &a, double2 X, Visit visit, Stop stop)
{
    int bx = (int)floor((X.x - a.ox) / a.h), by = (int)floor((X.y - a.oy) / a.h);
    bx = bx < 0 ? 0 : (bx >= a.BX ? a.BX - 1 : bx);
    by = by < 0 ? 0 : (by >= a.BY ? a.BY - 1 : by);
    const int rmax = max(max(bx, a.BX - 1 - bx), max(by, a.BY - 1 - by));
    auto bucket = [&](int i, int j) {
        if (i < 0 || i >= a.BX || j < 0 || j >= a.BY) return;
        const int b = j * a.BX + i;
        for (int t = a.bstart[b]; t < a.bstart[b + 1]; ++t) {
            const int p = a.bnode[t];
            if (!a.bad[p]) visit(p);
        }
    };
    for (int r = 0; r <= rmax; ++r) {
        if (r > 0 && stop((double)(r - 1) * a.h)) break;
        if (r == 0) {
            bucket(bx, by);
            continue;
        }
        for (int i = bx - r; i <= bx + r; ++i) { bucket(i, by - r); bucket(i, by + r); }
        for (int j = by - r + 1; j <= by + r - 1; ++j) { bucket(bx - r, j); bucket(bx + r, j); }
    }
}

__device__ void plane_at(const NatArgs &a, int ia, int ib, int ic, double2 Q, double2 &u,
                         double4 &f)
{
    const double2 A = a.xy[ia], Bp = a.xy[ib], Cp = a.xy[ic];
    const double bx = Bp.x - A.x, by = Bp.y - A.y, cx = Cp.x - A.x, cy = Cp.y - A.y;
    const double qx = Q.x - A.x, qy = Q.y - A.y;
    const double den = bx * cy - by * cx;
    const double al = (qx * cy - qy * cx) / den, be = (bx * qy - by * qx) / den;
    const double2 ua = a.U[ia], ub = a.U[ib], uc = a.U[ic];
    u = make_double2(ua.x + al * (ub.x - ua.x) + be * (uc.x - ua.x),
                     ua.y + al * (ub.y - ua.y) + be * (uc.y - ua.y));
    if (a.F) {
        const double4 fa = a.F[ia], fb = a.F[ib], fc = a.F[ic];
        f = make_double4(fa.x + al * (fb.x - fa.x) + be * (fc.x - fa.x),
                         fa.y + al * (fb.y - fa.y) + be * (fc.y - fa.y),
                         fa.z + al * (fb.z - fa.z) + be * (fc.z - fa.z),
                         fa.w + al * (fb.w - fa.w) + be * (fc.w - fa.w));
    }
}

// Circle through a and c (edge vector e, length L, midpoint m, inner unit normal n) and
// node p strictly on the inner side: its centre is m + s n.
__device__ __forceinline__ bool bulge(double2 Pa, double ex, double ey, double L, double2 M,
                                      double nxv, double nyv, double2 P, double &s)
{
    const double px = P.x - Pa.x, py = P.y - Pa.y;
    const double cr = ex * py - ey * px;
    if (!(cr > 1e-12 * L * sqrt(px * px + py * py))) return false;
    const double dx = P.x - M.x, dy = P.y - M.y;
    s = (dx * dx + dy * dy - 0.25 * L * L) / (2.0 * (dx * nxv + dy * nyv));
    return true;
}

// One wave per hull-list node: the lanes share the search for the nearest hull edge,
// the rest runs on every lane with the same values and lane 0 writes.
__global__ __launch_bounds__(64) void k_nat_hullfill(NatArgs a)
{
    const int nh = a.ctl[3], h = a.ctl[0], ng = a.ctl[1];
    const int lane = threadIdx.x;
    for (int li = blockIdx.x; li < nh; li += gridDim.x) {
        const int q = a.lhull[li];
        const double2 Q = a.xy[q];
        double2 u = make_double2(NAN, NAN);
        double4 f = make_double4(NAN, NAN, NAN, NAN);
        if (ng == 0) {
            // no good node: NaN
        } else if (h < 3 || ng < 3) {
            // nearest good node, ties to the lowest index
            double bd = INFINITY;
            int bi = 0x7fffffff;
            ring_scan(a, Q,
                      [&](int p) {
                          const double2 P = a.xy[p];
                          const double d = (P.x - Q.x) * (P.x - Q.x) + (P.y - Q.y) * (P.y - Q.y);
                          if (d < bd || (d == bd && p < bi)) { bd = d; bi = p; }
                      },
                      [&](double D) { return bi != 0x7fffffff && D * D >= bd; });
            u = a.U[bi];
            if (a.F) f = a.F[bi];
        } else {
            // the hull point nearest q: least distance, ties to the lowest edge
            int be = 0x7fffffff;
            double bt = 0.0, bd = INFINITY;
            for (int e = lane; e < h; e += 64) {
                const double2 A = a.xy[a.hull[e]], C = a.xy[a.hull[e + 1 < h ? e + 1 : 0]];
                const double ex = C.x - A.x, ey = C.y - A.y;
                double tt = ((Q.x - A.x) * ex + (Q.y - A.y) * ey) / (ex * ex + ey * ey);
                tt = tt < 0.0 ? 0.0 : (tt > 1.0 ? 1.0 : tt);
                const double dx = A.x + tt * ex - Q.x, dy = A.y + tt * ey - Q.y;
                const double d = dx * dx + dy * dy;
                if (d < bd) { bd = d; be = e; bt = tt; }
            }
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) {
                const double od = __shfl_xor(bd, m, kWave), ot = __shfl_xor(bt, m, kWave);
                const int oe = __shfl_xor(be, m, kWave);
                if (od < bd || (od == bd && oe < be)) { bd = od; be = oe; bt = ot; }
            }
            if (bt > 0.0 && bt < 1.0) {
                // inside the hull edge A-C: good nodes a (largest s <= bt) and c (smallest
                // s > bt) on it next to the nearest point
                const int A = a.hull[be], C = a.hull[be + 1 < h ? be + 1 : 0];
                const double2 XA = a.xy[A], XC = a.xy[C];
                const double LAC = sqrt((XC.x - XA.x) * (XC.x - XA.x) + (XC.y - XA.y) * (XC.y - XA.y));
                const double2 Bp = make_double2(XA.x + bt * (XC.x - XA.x), XA.y + bt * (XC.y - XA.y));
                double sa = 0.0, sc = 1.0;
                int ia = A, ic = C;
                ring_scan(a, Bp,
                          [&](int p) {
                              double sp;
                              if (!on_seg(XA, XC, a.xy[p], sp)) return;
                              if (sp <= bt && (sp > sa || (sp == sa && p < ia))) { sa = sp; ia = p; }
                              if (sp > bt && (sp < sc || (sp == sc && p < ic))) { sc = sp; ic = p; }
                          },
                          [&](double D) { return D >= (bt - sa) * LAC && D >= (sc - bt) * LAC; });
                // Delaunay third vertex of a-c: least bulge s over the inner side, then the
                // lowest index within the tolerance of it
                const double2 Pa = a.xy[ia], Pc = a.xy[ic];
                const double ex = Pc.x - Pa.x, ey = Pc.y - Pa.y;
                const double L = sqrt(ex * ex + ey * ey);
                const double2 M = make_double2(Pa.x + 0.5 * ex, Pa.y + 0.5 * ey);
                const double nxv = -ey / L, nyv = ex / L;
                double smin = INFINITY;
                // nodes farther than D from m bulge more than (D^2 - L^2/4) / (2 D)
                auto beyond = [&](double D, double lim) {
                    return D > 0.5 * L && (D * D - 0.25 * L * L) / (2.0 * D) > lim;
                };
                ring_scan(a, M,
                          [&](int p) {
                              double sp;
                              if (bulge(Pa, ex, ey, L, M, nxv, nyv, a.xy[p], sp) && sp < smin) smin = sp;
                          },
                          [&](double D) {
                              return smin < INFINITY && beyond(D, smin + 1e-12 * (fabs(smin) + L));
                          });
                const double lim = smin + 1e-12 * (fabs(smin) + L);
                int ir = 0x7fffffff;
                ring_scan(a, M,
                          [&](int p) {
                              double sp;
                              if (p < ir && bulge(Pa, ex, ey, L, M, nxv, nyv, a.xy[p], sp) && sp <= lim)
                                  ir = p;
                          },
                          [&](double D) { return beyond(D, lim); });
                if (ir != 0x7fffffff) plane_at(a, ia, ic, ir, Q, u, f);
                else if (lane == 0) nat_error(a);
            } else {
                // at the hull vertex v: the good nodes next to v along its two hull edges
                const int vi = bt <= 0.0 ? be : (be + 1 < h ? be + 1 : 0);
                const int v = a.hull[vi], Pv = a.hull[vi > 0 ? vi - 1 : h - 1],
                          Nv = a.hull[vi + 1 < h ? vi + 1 : 0];
                const double2 XP = a.xy[Pv], XV = a.xy[v], XN = a.xy[Nv];
                const double L1 = sqrt((XV.x - XP.x) * (XV.x - XP.x) + (XV.y - XP.y) * (XV.y - XP.y));
                const double L2 = sqrt((XN.x - XV.x) * (XN.x - XV.x) + (XN.y - XV.y) * (XN.y - XV.y));
                double s1 = 0.0, s2 = 1.0;
                int i1 = Pv, i2 = Nv;
                ring_scan(a, XV,
                          [&](int p) {
                              if (p == v) return;
                              double sp;
                              const double2 X = a.xy[p];
                              if (on_seg(XP, XV, X, sp) && sp < 1.0 && (sp > s1 || (sp == s1 && p < i1))) {
                                  s1 = sp; i1 = p;
                              }
                              if (on_seg(XV, XN, X, sp) && sp > 0.0 && (sp < s2 || (sp == s2 && p < i2))) {
                                  s2 = sp; i2 = p;
                              }
                          },
                          [&](double D) { return D >= (1.0 - s1) * L1 && D >= s2 * L2; });
                plane_at(a, v, i2, i1, Q, u, f);
            }
        }
        if (lane == 0) {
            a.U[q] = u;
            if (a.F) a.F[q] = f;
        }
    }
}

__global__ void k_nat_clear(int *err) { *err = 0; }

}  // namespace

static hipError_t natgrid_setup_impl(NatGrid &g, const double *coords, int nn, hipStream_t st)
{
    g.nn = nn;
    if (nn <= 0) return hipSuccess;
    double xmin = coords[0], xmax = coords[0], ymin = coords[1], ymax = coords[1];
    for (int i = 1; i < nn; ++i) {
        xmin = std::min(xmin, coords[2 * i]); xmax = std::max(xmax, coords[2 * i]);
        ymin = std::min(ymin, coords[2 * i + 1]); ymax = std::max(ymax, coords[2 * i + 1]);
    }
    if (!std::isfinite(xmin) || !std::isfinite(xmax) || !std::isfinite(ymin) || !std::isfinite(ymax))
        return hipErrorInvalidValue;
    // bucket side ~ the mean node spacing (one node per bucket on a uniform grid), and at
    // least extent / (2 sqrt(nn) + 1), so a thin or collinear set still gets at most
    // (2 sqrt(nn) + 3)^2 buckets
    const double ext = std::max(xmax - xmin, ymax - ymin);
    const double rn = std::sqrt((double)nn);
    double h = std::sqrt((xmax - xmin) * (ymax - ymin) / nn);
    h = std::max(h, ext / (2.0 * rn + 1.0));
    // a uniform lattice: its spacing (node i+1 to the right of node i)
    if (nn >= 2 && coords[3] == coords[1] && coords[2] > coords[0]) {
        const double sp = coords[2] - coords[0];
        if (std::fabs(sp - h) < 0.5 * h) h = sp;
    }
    if (!(h > 0)) h = 1.0;
    g.h = h;
    g.ox = xmin - 0.5 * h;
    g.oy = ymin - 0.5 * h;
    g.BX = (int)std::floor((xmax - g.ox) / h) + 1;
    g.BY = (int)std::floor((ymax - g.oy) / h) + 1;
    const size_t nb = (size_t)g.BX * g.BY;
    std::vector<int> cnt(nb + 1, 0), bnode(nn), rnode(nn);
    std::vector<int> bk(nn);
    for (int i = 0; i < nn; ++i) {
        int bx = (int)std::floor((coords[2 * i] - g.ox) / h);
        int by = (int)std::floor((coords[2 * i + 1] - g.oy) / h);
        bx = std::min(std::max(bx, 0), g.BX - 1);
        by = std::min(std::max(by, 0), g.BY - 1);
        bk[i] = by * g.BX + bx;
        cnt[bk[i] + 1]++;
    }
    for (size_t b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < nn; ++i) bnode[pos[bk[i]]++] = i;
    // rows: equal y, x ascending
    for (int i = 0; i < nn; ++i) rnode[i] = i;
    std::sort(rnode.begin(), rnode.end(), [&](int p, int q) {
        if (coords[2 * p + 1] != coords[2 * q + 1]) return coords[2 * p + 1] < coords[2 * q + 1];
        if (coords[2 * p] != coords[2 * q]) return coords[2 * p] < coords[2 * q];
        return p < q;
    });
    std::vector<int> rstart;
    for (int k = 0; k < nn; ++k)
        if (k == 0 || coords[2 * rnode[k] + 1] != coords[2 * rnode[k - 1] + 1]) rstart.push_back(k);
    rstart.push_back(nn);
    g.R = (int)rstart.size() - 1;
    hipError_t e;
    if ((e = g.bstart.ensure(sizeof(int) * (nb + 1))) || (e = g.bnode.ensure(sizeof(int) * nn)) ||
        (e = g.rstart.ensure(sizeof(int) * (g.R + 1))) || (e = g.rnode.ensure(sizeof(int) * nn)) ||
        (e = g.work.ensure(sizeof(int) * (3 * (size_t)g.R + 6 * (size_t)g.R + 4 +
                                          2 * (size_t)g.R + 2 + 8 + 4 * (size_t)nn))) ||
        (e = g.big.ensure((sizeof(int) + 2 * sizeof(double)) * 2 * (size_t)kNatBig * kNatBigThreads)))
        return e;
    if ((e = hipMemcpyAsync(g.bstart.p, cnt.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(g.bnode.p, bnode.data(), sizeof(int) * nn, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(g.rstart.p, rstart.data(), sizeof(int) * (g.R + 1), hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(g.rnode.p, rnode.data(), sizeof(int) * nn, hipMemcpyHostToDevice, st)))
        return e;
    return hipStreamSynchronize(st);  // the host vectors go out of scope
}

hipError_t natgrid_setup(NatGrid &g, const double *coords, int nn, hipStream_t st)
{
    try {
        return natgrid_setup_impl(g, coords, nn, st);
    } catch (...) {  // host allocation: no exception crosses the C ABI
        return hipErrorOutOfMemory;
    }
}

hipError_t natural_err_clear(int *err, hipStream_t st)
{
    hipLaunchKernelGGL(k_nat_clear, dim3(1), dim3(1), 0, st, err);
    return hipGetLastError();
}

hipError_t natural_fill(const NatGrid &g, const double *coords, const int *bad, double *U,
                        double *F, hipStream_t st, int *err)
{
    if (g.nn <= 0) return hipSuccess;
    NatArgs a{};
    a.xy = reinterpret_cast<const double2 *>(coords);
    a.bad = bad;
    a.U = reinterpret_cast<double2 *>(U);
    a.F = reinterpret_cast<double4 *>(F);
    a.nn = g.nn;
    a.BX = g.BX;
    a.BY = g.BY;
    a.ox = g.ox;
    a.oy = g.oy;
    a.h = g.h;
    a.bstart = g.bstart.as<int>();
    a.bnode = g.bnode.as<int>();
    a.R = g.R;
    a.rstart = g.rstart.as<int>();
    a.rnode = g.rnode.as<int>();
    int *w = g.work.as<int>();
    a.rows = w;
    a.chain = a.rows + 3 * (size_t)g.R;
    a.hull = a.chain + 6 * (size_t)g.R + 4;
    a.ctl = a.hull + 2 * (size_t)g.R + 2;
    a.lin = a.ctl + 8;
    a.lhull = a.lin + g.nn;
    a.lbig = a.lhull + g.nn;
    a.lbig2 = a.lbig + g.nn;
    a.err = err;
    const size_t nbig = 2 * (size_t)kNatBig * kNatBigThreads;
    a.bvx = g.big.as<double>();
    a.bvy = a.bvx + nbig;
    a.btag = reinterpret_cast<int *>(a.bvy + nbig);
    hipLaunchKernelGGL(k_nat_rows, dim3((g.R + 3) / 4), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_nat_hull, dim3(1), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_nat_classify, dim3((g.nn + 255) / 256), dim3(256), 0, st, a);
    const int nbs = std::min((g.nn + kNatThreads - 1) / kNatThreads, 4096);
    hipLaunchKernelGGL(k_nat_sibson, dim3(nbs), dim3(kNatThreads), 0, st, a);
    hipLaunchKernelGGL(k_nat_sibson_wave, dim3(std::min(g.nn, 2048)), dim3(64), 0, st, a);
    hipLaunchKernelGGL(k_nat_sibson_big, dim3(kNatBigThreads / 64), dim3(64), 0, st, a);
    const int nbh = std::min(g.nn, 2048);
    hipLaunchKernelGGL(k_nat_hullfill, dim3(nbh), dim3(64), 0, st, a);
    return hipGetLastError();
}

int natural_errors(const NatGrid &g, hipStream_t st)
{
    int ctl[5] = {0, 0, 0, 0, 0};
    const int *w = g.work.as<int>();
    if (!w) return 0;
    const int *c = w + 3 * (size_t)g.R + 6 * (size_t)g.R + 4 + 2 * (size_t)g.R + 2;
    if (hipMemcpyAsync(ctl, c, sizeof ctl, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
    return ctl[4];
}

}  // namespace aldic
