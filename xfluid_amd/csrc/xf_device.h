// xf_device.h -- shared device helpers for the gfx950 substep kernels.
//
// All parity kernels are compiled with -ffp-contract=off: every expression
// below is evaluated in the operand order of the reference expression it
// restates, with no fused multiply-add, so per-element results are bit-exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define XF_AIR 0
#define XF_FLUID 1
// pressure operators (the values of XF_PRESSURE_VARIATIONAL / XF_PRESSURE_NORMAL)
#define OP_VARIATIONAL 0
#define OP_NORMAL 1
#define XF_SOLID 2

// Grid geometry passed by value to every kernel.
struct Geo {
    int ni, nj, nk;
    long N;      // ni*nj*nk
    long nn;     // ni*nj
    double dx;
};

__host__ __device__ inline long cidx(const Geo& g, int i, int j, int k) { return i + (long)g.ni * j + g.nn * k; }
__host__ __device__ inline bool cin(const Geo& g, int i, int j, int k) {
    return i >= 0 && j >= 0 && k >= 0 && i < g.ni && j < g.nj && k < g.nk;
}

// xfluidutil.cpp:3-21
__device__ inline int pos2idx(double p, double dx) { return (int)(p / dx); }
__device__ inline double idx2pos(int i, double dx) { return (i + 0.5) * dx; }

// xfluidutil.cpp:53-65
__device__ inline double trilerp8(const double* p, double x, double y, double z) {
    double v0 = (1.0 - x) * p[0] + x * p[1];
    double v1 = (1.0 - x) * p[2] + x * p[3];
    double v2 = (1.0 - x) * p[4] + x * p[5];
    double v3 = (1.0 - x) * p[6] + x * p[7];
    v0 = (1.0 - y) * v0 + y * v1;
    v1 = (1.0 - y) * v2 + y * v3;
    return (1.0 - z) * v0 + z * v1;
}

// Wyvill "soft object" kernel, FluidSimulator.cpp:917-932 and
// directeddistancefield.cpp:27-37,246-256.
struct Wyvill {
    double r2, c0, c1, c2;
};
__host__ __device__ inline Wyvill make_wyvill(double r) {
    Wyvill w;
    w.r2 = r * r;
    w.c0 = (4.0 / 9.0) * (1.0 / (w.r2 * w.r2 * w.r2));
    w.c1 = (17.0 / 9.0) * (1.0 / (w.r2 * w.r2));
    w.c2 = (22.0 / 9.0) * (1.0 / (w.r2));
    return w;
}
__device__ inline double wyvill(const Wyvill& w, double s) {
    return s < w.r2 ? 1.0 - w.c0 * s * s * s + w.c1 * s * s - w.c2 * s : 0.0;
}

// Stencil code of a pressure row: 2 bits per neighbour direction giving the
// off-diagonal coefficient class (0: none, 1: -0.5, 2: -1.0), plus the cell's
// material in bits 12-13.  Direction order of the bit fields:
//   0: -k (c-nn), 1: -j (c-ni), 2: -i (c-1), 3: +i (c+1), 4: +j (c+ni), 5: +k (c+nn)
// i.e. the CSR column order of sparse_matrix.h:266-276 with the diagonal taken out.
enum { D_ZM = 0, D_YM = 1, D_XM = 2, D_XP = 3, D_YP = 4, D_ZP = 5 };
__host__ __device__ inline int code_dir(uint16_t c, int d) { return (c >> (2 * d)) & 3; }
__host__ __device__ inline int code_mat(uint16_t c) { return (c >> 12) & 3; }
__host__ __device__ inline double code_coef(int cls) { return cls == 2 ? -1.0 : (cls == 1 ? -0.5 : 0.0); }
// Diagonal: fluid rows sum fluid 1 / air 2 / solid 0.5 over the six faces,
// solid rows 0.5 per fluid neighbour (variational_pressuresolver.cpp:294-321).
// All terms are exact halves, so the order of the sum does not matter.
// Counted per class with bit operations on the six 2-bit fields (class 2 in
// n2 fields, class 1 in n1; classes 0 and 3 count 2.0 on a fluid row).
// A fluid row of the NormalPressureSolver operator carries CODE_NORMAL_ROW in
// the material field: its diagonal counts the in-range non-solid neighbours,
// fluid (class 2, off-diagonal -1) and air (class 0); solid and outside
// neighbours are class 3 (pressuresolver.cpp:241-272).
#define CODE_NORMAL_ROW 3
__host__ __device__ inline double code_diag(uint16_t c) {
    const int m = code_mat(c);
    const unsigned lo = c & 0x555u, hi = (c >> 1) & 0x555u;
    const int n2 = __builtin_popcount(hi & ~lo), n1 = __builtin_popcount(lo & ~hi);
    if (m == XF_FLUID) return n2 * 1.0 + n1 * 0.5 + (6 - n2 - n1) * 2.0;
    if (m == XF_SOLID) return n1 * 0.5;
    if (m == CODE_NORMAL_ROW) return (double)(n2 + __builtin_popcount(~(lo | hi) & 0x555u));
    return 0.0;
}

// Workgroups go to the 8 XCDs of the MI355X round-robin (block b to XCD b % 8),
// and each XCD has its own L2.  For gathers whose neighbouring blocks read the
// same data, give each XCD a contiguous run of logical blocks instead of every
// 8th one (a bijection on [0, nb): the last nb % 8 blocks keep their index).
__device__ inline long xcd_block(long b, long nb) {
    const long full = nb & ~7l;
    if (b >= full) return b;
    return (b & 7) * (full >> 3) + (b >> 3);
}

template <class T>
__device__ inline T warp_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
template <class T>
__device__ inline T warp_max(T v) {
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_down(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Block reduction (blockDim multiple of 64, <= 1024) with a fixed tree order,
// so the result is deterministic for a fixed launch shape.
template <class T, bool MAX>
__device__ inline T block_reduce(T v, T* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = MAX ? warp_max(v) : warp_sum(v);
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (wid == 0) {
        v = lane < nw ? sh[lane] : (T)0;
        v = MAX ? warp_max(v) : warp_sum(v);
    }
    __syncthreads();
    return v;  // valid in thread 0
}

// (i, j, k) of flat index t of an si x sj x * array, in 32-bit arithmetic
// (grids hold fewer than 2^30 cells, so every cell and face index fits)
__device__ inline void unflat3(long t, int si, int sj, int* i, int* j, int* k) {
    const unsigned u = (unsigned)t, r = u / (unsigned)si;
    *i = (int)(u - r * (unsigned)si);
    *j = (int)(r % (unsigned)sj);
    *k = (int)(r / (unsigned)sj);
}

// material of cell (i,j,k), `outside` for cells beyond the grid
__device__ inline uint8_t mat_at(const Geo& g, const uint8_t* mat, int i, int j, int k, uint8_t outside) {
    return cin(g, i, j, k) ? mat[cidx(g, i, j, k)] : outside;
}

// ---------------------------------------------------------------------------
// _getVelocity (FluidSimulator.cpp:814-894): staggered trilinear sample with
// the reference's clamps; outside the box the velocity is 0.
struct Faces {
    const double *u, *v, *w;
};
__device__ inline void sample(const Geo& g, const Faces& F, const double* p, double* out) {
    const double dx = g.dx;
    if (!((p[0] >= 0.0) && (p[1] >= 0.0) && (p[2] >= 0.0) && (p[0] <= dx * g.ni) && (p[1] <= dx * g.nj) &&
          (p[2] <= dx * g.nk))) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double invdx = 1.0 / dx;
    double t[8], x, y, z;
    int i, j, k;
    x = p[0] * invdx; y = (p[1] - 0.5 * dx) * invdx; z = (p[2] - 0.5 * dx) * invdx;
    i = (int)x;
    j = y < 0.0 ? 0 : (int)y; j = j > (g.nj - 2) ? (g.nj - 2) : j;
    k = z < 0.0 ? 0 : (int)z; k = k > (g.nk - 2) ? (g.nk - 2) : k;
    x -= i; y -= j; z -= k;
    {
        const long si = g.ni + 1, sij = si * g.nj;
        const double* b = F.u + i + si * j + sij * k;
        t[0] = b[0]; t[1] = b[1]; t[2] = b[si]; t[3] = b[si + 1];
        t[4] = b[sij]; t[5] = b[sij + 1]; t[6] = b[sij + si]; t[7] = b[sij + si + 1];
    }
    out[0] = trilerp8(t, x, y, z);
    x = (p[0] - 0.5 * dx) * invdx; y = p[1] * invdx; z = (p[2] - 0.5 * dx) * invdx;
    j = (int)y;
    i = x < 0.0 ? 0 : (int)x; i = i > (g.ni - 2) ? (g.ni - 2) : i;
    k = z < 0.0 ? 0 : (int)z; k = k > (g.nk - 2) ? (g.nk - 2) : k;
    x -= i; y -= j; z -= k;
    {
        const long si = g.ni, sij = si * (g.nj + 1);
        const double* b = F.v + i + si * j + sij * k;
        t[0] = b[0]; t[1] = b[1]; t[2] = b[si]; t[3] = b[si + 1];
        t[4] = b[sij]; t[5] = b[sij + 1]; t[6] = b[sij + si]; t[7] = b[sij + si + 1];
    }
    out[1] = trilerp8(t, x, y, z);
    z = p[2] * invdx; y = (p[1] - 0.5 * dx) * invdx; x = (p[0] - 0.5 * dx) * invdx;
    k = (int)z;
    j = y < 0.0 ? 0 : (int)y; j = j > (g.nj - 2) ? (g.nj - 2) : j;
    i = x < 0.0 ? 0 : (int)x; i = i > (g.ni - 2) ? (g.ni - 2) : i;
    x -= i; y -= j; z -= k;
    {
        const long si = g.ni, sij = si * g.nj;
        const double* b = F.w + i + si * j + sij * k;
        t[0] = b[0]; t[1] = b[1]; t[2] = b[si]; t[3] = b[si + 1];
        t[4] = b[sij]; t[5] = b[sij + 1]; t[6] = b[sij + si]; t[7] = b[sij + si + 1];
    }
    out[2] = trilerp8(t, x, y, z);
}

// The cells of the active rows' bounding box (i0..i1, j0..j1, k0..k1), walked
// by a grid-stride loop.  Outside the box every PCG vector is zero.
struct BoxRange {
    long base, ni, nn;                // first cell (i0, j0, k0) and the grid's strides
    unsigned iw, jw, n, si, sj, sk;   // box extents in i and j, its cells, the stride in (i, j, k) digits
};
__host__ __device__ inline BoxRange box_range(long ni, long nj, long nk, const int box[6], unsigned stride) {
    (void)nk;
    BoxRange r;
    r.ni = ni;
    r.nn = ni * nj;
    r.base = (long)box[4] * r.nn + (long)box[2] * ni + box[0];
    r.iw = (unsigned)(box[1] - box[0] + 1);
    r.jw = (unsigned)(box[3] - box[2] + 1);
    r.n = r.iw * r.jw * (unsigned)(box[5] - box[4] + 1);
    r.si = stride % r.iw;
    r.sj = (stride / r.iw) % r.jw;
    r.sk = stride / (r.iw * r.jw);
    return r;
}
// The walk over the box's cells in (k, j, i) order: sequence number t is the
// mixed-radix number (kq, jq, iq); a grid-stride step adds the stride's digits
// with carries (no division per cell).
struct BoxWalk { unsigned kq, jq, iq; };
__device__ inline void bw_init(const BoxRange& b, unsigned t, BoxWalk& w) {
    w.iq = t % b.iw;
    const unsigned q = t / b.iw;
    w.jq = q % b.jw;
    w.kq = q / b.jw;
}
__device__ inline void bw_next(const BoxRange& b, BoxWalk& w) {
    w.iq += b.si;
    w.jq += b.sj;
    w.kq += b.sk;
    if (w.iq >= b.iw) { w.iq -= b.iw; ++w.jq; }
    if (w.jq >= b.jw) { w.jq -= b.jw; ++w.kq; }
}
__device__ inline long bw_cell(const BoxRange& b, const BoxWalk& w) {
    return b.base + (long)w.kq * b.nn + (long)w.jq * b.ni + w.iq;
}

// collision.cpp:293-385 (voxel walk to the first solid cell)
__device__ inline bool voxel_hit(const Geo& g, const uint8_t* mat, const double* a, const double* b, int* vox) {
    const double invdx = 1.0 / g.dx;
    double p0[3] = {a[0] * invdx, a[1] * invdx, a[2] * invdx};
    double p1[3] = {b[0] * invdx, b[1] * invdx, b[2] * invdx};
    int g0[3], g1[3], sd[3], gg[3];
    double gp[3], vv[3];
    for (int t = 0; t < 3; ++t) {
        g0[t] = (int)floor(p0[t]); g1[t] = (int)floor(p1[t]);
        sd[t] = g1[t] > g0[t] ? 1 : g1[t] < g0[t] ? -1 : 0;
        gg[t] = g0[t];
        gp[t] = g0[t] + (g1[t] > g0[t] ? 1 : 0);
        vv[t] = p1[t] == p0[t] ? 1 : p1[t] - p0[t];
    }
    double vxvy = vv[0] * vv[1], vxvz = vv[0] * vv[2], vyvz = vv[1] * vv[2];
    double errx = (gp[0] - p0[0]) * vyvz, erry = (gp[1] - p0[1]) * vxvz, errz = (gp[2] - p0[2]) * vxvy;
    double derrx = sd[0] * vyvz, derry = sd[1] * vxvz, derrz = sd[2] * vxvy;
    // the walk ends after |dg| steps; bound it for safety
    int budget = abs(g1[0] - g0[0]) + abs(g1[1] - g0[1]) + abs(g1[2] - g0[2]) + 1;
    for (int step = 0; step <= budget; ++step) {
        if (cin(g, gg[0], gg[1], gg[2]) && mat[cidx(g, gg[0], gg[1], gg[2])] == XF_SOLID) {
            vox[0] = gg[0]; vox[1] = gg[1]; vox[2] = gg[2];
            return true;
        }
        if (gg[0] == g1[0] && gg[1] == g1[1] && gg[2] == g1[2]) break;
        double xr = fabs(errx), yr = fabs(erry), zr = fabs(errz);
        if (sd[0] != 0 && (sd[1] == 0 || xr < yr) && (sd[2] == 0 || xr < zr)) { gg[0] += sd[0]; errx += derrx; }
        else if (sd[1] != 0 && (sd[2] == 0 || yr < zr)) { gg[1] += sd[1]; erry += derry; }
        else if (sd[2] != 0) { gg[2] += sd[2]; errz += derrz; }
    }
    return false;
}

// collision.cpp:396-439, including its repeated dir.x guard.  Also the axis
// whose slab the ray enters last (the face it hits) and whether that face is
// the box's upper bound.
__device__ inline bool ray_box(const double* p0, const double* din, const double* bmin, const double* bmax,
                               double* hit, int* axis = nullptr, bool* upper = nullptr) {
    double dir[3] = {din[0], din[1], din[2]};
    const double eps = 1e-10;
    if (fabs(dir[0]) < eps) dir[0] = dir[0] < 0 ? -eps : eps;
    if (fabs(dir[1]) < eps) dir[1] = dir[1] < 0 ? -eps : eps;
    if (fabs(dir[0]) < eps) dir[2] = dir[2] < 0 ? -eps : eps;
    double inv[3] = {1.0 / dir[0], 1.0 / dir[1], 1.0 / dir[2]};
    double t1 = (bmin[0] - p0[0]) * inv[0], t2 = (bmax[0] - p0[0]) * inv[0];
    double tmin = fmin(t1, t2), tmax = fmax(t1, t2);
    const double m0 = tmin;
    t1 = (bmin[1] - p0[1]) * inv[1]; t2 = (bmax[1] - p0[1]) * inv[1];
    const double m1 = fmin(t1, t2);
    tmin = fmax(tmin, m1); tmax = fmin(tmax, fmax(t1, t2));
    t1 = (bmin[2] - p0[2]) * inv[2]; t2 = (bmax[2] - p0[2]) * inv[2];
    const double m2 = fmin(t1, t2);
    tmin = fmax(tmin, m2); tmax = fmin(tmax, fmax(t1, t2));
    if (tmax > fmax(tmin, 0.0)) {
        for (int a = 0; a < 3; ++a) hit[a] = p0[a] + dir[a] * tmin;
        if (axis) {
            const int ax = tmin == m2 ? 2 : tmin == m1 ? 1 : (tmin == m0 ? 0 : 0);
            *axis = ax;
            *upper = inv[ax] < 0.0;
        }
        return true;
    }
    return false;
}

// The end of a particle's move in _advectParticles (FluidSimulator.cpp:686-722):
// the integrated target np_ when its cell is in range and not solid; else the
// first solid voxel on the segment pos -> np_ (Collision::getLineSegmentVoxel-
// Intersection), the ray's entry into that voxel's box (the reference passes
// the voxel centre as the box's min corner, aabb.cpp:11-13), pulled back by
// (float)(0.05 dx) along the ray, accepted if its cell is in range and not
// solid; otherwise the particle keeps pos.  `c` records what the adjoint needs.
enum { MOVE_TARGET = 0, MOVE_KEPT = 1, MOVE_PULLED = 2 };
struct Collision {
    double rn[3], l, tmin, back;
    int axis;
};
__device__ inline int resolve_move(const Geo& g, const uint8_t* mat, const double* pos, const double* np_,
                                   double* out, Collision* c) {
    int gi = pos2idx(np_[0], g.dx), gj = pos2idx(np_[1], g.dx), gk = pos2idx(np_[2], g.dx);
    if (cin(g, gi, gj, gk) && mat[cidx(g, gi, gj, gk)] != XF_SOLID) {
        out[0] = np_[0]; out[1] = np_[1]; out[2] = np_[2];
        return MOVE_TARGET;
    }
    int vox[3];
    if (!voxel_hit(g, mat, pos, np_, vox)) return MOVE_KEPT;
    double dv[3] = {np_[0] - pos[0], np_[1] - pos[1], np_[2] - pos[2]};
    double l = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    double rn[3] = {dv[0] / l, dv[1] / l, dv[2] / l};
    double bmin[3] = {idx2pos(vox[0], g.dx), idx2pos(vox[1], g.dx), idx2pos(vox[2], g.dx)};
    double bmax[3] = {bmin[0] + g.dx, bmin[1] + g.dx, bmin[2] + g.dx};
    double cp[3];
    int axis = 0;
    bool upper = false;
    if (!ray_box(pos, rn, bmin, bmax, cp, &axis, &upper)) return MOVE_KEPT;
    const double back = (double)(float)(0.05 * g.dx);
    double res[3] = {cp[0] - rn[0] * back, cp[1] - rn[1] * back, cp[2] - rn[2] * back};
    int ri = pos2idx(res[0], g.dx), rj = pos2idx(res[1], g.dx), rk = pos2idx(res[2], g.dx);
    if (!(cin(g, ri, rj, rk) && mat[cidx(g, ri, rj, rk)] != XF_SOLID)) return MOVE_KEPT;
    out[0] = res[0]; out[1] = res[1]; out[2] = res[2];
    for (int a = 0; a < 3; ++a) c->rn[a] = rn[a];
    c->l = l;
    c->axis = axis;
    c->tmin = ((upper ? bmax[axis] : bmin[axis]) - pos[axis]) / rn[axis];
    c->back = back;
    return MOVE_PULLED;
}
