// xf_grid.hip -- per-cell and per-face stage kernels of the substep.
//
// Each kernel is a gather (one thread per output cell or face), so results are
// deterministic and, compiled with -ffp-contract=off, bit-identical to the
// serial reference loops they restate.
#include "xf_internal.h"

// ---------------------------------------------------------------------------
// _updateMaterial (FluidSimulator.cpp:402-427): every non-solid cell becomes
// air, then any cell holding a particle becomes fluid.  The particle count per
// cell comes from the cell sort, so no atomics touch the material.
__global__ void k_material(Geo g, uint8_t* mat, const int32_t* cnt) {
    long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    uint8_t m = mat[c];
    if (m != XF_SOLID) mat[c] = cnt[c] > 0 ? XF_FLUID : XF_AIR;
}
void launch_material(const Geo& g, uint8_t* mat, const int32_t* cell_count, hipStream_t st) {
    k_material<<<xf_blocks(g.N), XF_BLOCK, 0, st>>>(g, mat, cell_count);
}

// ---------------------------------------------------------------------------
// _applyForces (FluidSimulator.cpp:429-499): faces bordering a fluid cell get
// += a*dt along the axis.
__global__ void k_forces(Geo g, const uint8_t* mat, double* f, int axis, double dv) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1), sk = g.nk + (axis == 2);
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    long nf = (long)si * sj * sk;
    if (t >= nf) return;
    int i, j, k;
    unflat3(t, si, sj, &i, &j, &k);
    int oi = axis == 0, oj = axis == 1, ok = axis == 2;
    bool a = mat_at(g, mat, i, j, k, XF_AIR) == XF_FLUID;
    bool b = mat_at(g, mat, i - oi, j - oj, k - ok, XF_AIR) == XF_FLUID;
    if (a || b) f[t] += dv;
}
void launch_forces(const Geo& g, const uint8_t* mat, double* f, int axis, double dv, hipStream_t st) {
    long nf = (long)(g.ni + (axis == 0)) * (g.nj + (axis == 1)) * (g.nk + (axis == 2));
    k_forces<<<xf_blocks(nf), XF_BLOCK, 0, st>>>(g, mat, f, axis, dv);
}

// ---------------------------------------------------------------------------
// Stencil codes + right-hand side, variational_pressuresolver.cpp:231-326.
// The reference visits fluid cells with i slowest and scatters into its own
// row and into solid neighbours' rows.  As a gather:
//  * a fluid row sums its six face terms in neighbour order +i,-i,+j,-j,+k,-k;
//  * a solid row receives one term per fluid neighbour, in the order the
//    reference visits those neighbours (i-major loop): -i,-j,-k,+k,+j,+i.
// Direction classes stored in the code (2 bits each, see xf_device.h):
//   fluid row: 0 air neighbour, 1 solid (-0.5), 2 fluid (-1), 3 outside grid;
//   solid row: 1 fluid neighbour (-0.5), else 0.
__device__ inline double face_u(const Geo& g, const double* u, int i, int j, int k) {
    return u[i + (long)(g.ni + 1) * (j + (long)g.nj * k)];
}
__device__ inline double face_v(const Geo& g, const double* v, int i, int j, int k) {
    return v[i + (long)g.ni * (j + (long)(g.nj + 1) * k)];
}
__device__ inline double face_w(const Geo& g, const double* w, int i, int j, int k) {
    return w[i + (long)g.ni * (j + (long)g.nj * k)];
}

__global__ void k_stencil_rhs(Geo g, const uint8_t* mat, const double* u, const double* v, const double* w,
                              uint16_t* code, double* b) {
    long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    int i, j, k;
    unflat3(c, g.ni, g.nj, &i, &j, &k);
    uint8_t m = mat[c];
    const uint8_t OUT = 3;
    // neighbour materials in code direction order: -k,-j,-i,+i,+j,+k
    uint8_t nb[6];
    nb[D_ZM] = mat_at(g, mat, i, j, k - 1, OUT);
    nb[D_YM] = mat_at(g, mat, i, j - 1, k, OUT);
    nb[D_XM] = mat_at(g, mat, i - 1, j, k, OUT);
    nb[D_XP] = mat_at(g, mat, i + 1, j, k, OUT);
    nb[D_YP] = mat_at(g, mat, i, j + 1, k, OUT);
    nb[D_ZP] = mat_at(g, mat, i, j, k + 1, OUT);
    // face velocity on each side
    double fc[6];
    fc[D_ZM] = face_w(g, w, i, j, k);
    fc[D_ZP] = face_w(g, w, i, j, k + 1);
    fc[D_YM] = face_v(g, v, i, j, k);
    fc[D_YP] = face_v(g, v, i, j + 1, k);
    fc[D_XM] = face_u(g, u, i, j, k);
    fc[D_XP] = face_u(g, u, i + 1, j, k);
    uint16_t cd = (uint16_t)((m & 3) << 12);
    double bb = 0;
    if (m == XF_FLUID) {
        for (int d = 0; d < 6; ++d) {
            int cl = nb[d] == XF_FLUID ? 2 : nb[d] == XF_SOLID ? 1 : nb[d] == OUT ? 3 : 0;
            cd |= (uint16_t)(cl << (2 * d));
        }
        // reference neighbour order +i,-i,+j,-j,+k,-k; vsign -1 on the + side
        const int ord[6] = {D_XP, D_XM, D_YP, D_YM, D_ZP, D_ZM};
        for (int e = 0; e < 6; ++e) {
            int d = ord[e];
            double vsign = (e & 1) ? 1.0 : -1.0;
            uint8_t nm = nb[d];
            if (nm == XF_FLUID || nm == XF_AIR) bb += vsign * fc[d];
            else bb += vsign * fc[d] * 0.5;
        }
    } else if (m == XF_SOLID) {
        for (int d = 0; d < 6; ++d)
            if (nb[d] == XF_FLUID) cd |= (uint16_t)(1 << (2 * d));
        // fluid neighbour f sees this cell on its opposite side: its vsign is
        // -1 when this cell is f's + neighbour (f on our - side), +1 otherwise
        const int ord[6] = {D_XM, D_YM, D_ZM, D_ZP, D_YP, D_XP};
        for (int e = 0; e < 6; ++e) {
            int d = ord[e];
            if (nb[d] != XF_FLUID) continue;
            double vsign_f = (d == D_XM || d == D_YM || d == D_ZM) ? -1.0 : 1.0;
            bb += -vsign_f * fc[d] * 0.5;
        }
    }
    code[c] = cd;
    b[c] = bb;
}
// NormalPressureSolver::_calculate_coefficient (pressuresolver.cpp:220-347):
// rows for fluid cells only; diagonal = in-range non-solid neighbours, -1 per
// fluid neighbour; b = +u(i) - u(i+1) + v(j) - v(j+1) + w(k) - w(k+1), in that
// order, each term only when the cell on that side is in range and not solid.
__global__ void k_stencil_rhs_normal(Geo g, const uint8_t* mat, const double* u, const double* v, const double* w,
                                     uint16_t* code, double* b) {
    long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    int i, j, k;
    unflat3(c, g.ni, g.nj, &i, &j, &k);
    const uint8_t m = mat[c];
    if (m != XF_FLUID) {  // an empty row
        code[c] = (uint16_t)((m & 3) << 12);
        b[c] = 0.0;
        return;
    }
    uint8_t nb[6];
    nb[D_ZM] = mat_at(g, mat, i, j, k - 1, XF_SOLID);
    nb[D_YM] = mat_at(g, mat, i, j - 1, k, XF_SOLID);
    nb[D_XM] = mat_at(g, mat, i - 1, j, k, XF_SOLID);
    nb[D_XP] = mat_at(g, mat, i + 1, j, k, XF_SOLID);
    nb[D_YP] = mat_at(g, mat, i, j + 1, k, XF_SOLID);
    nb[D_ZP] = mat_at(g, mat, i, j, k + 1, XF_SOLID);
    uint16_t cd = (uint16_t)(CODE_NORMAL_ROW << 12);
    for (int d = 0; d < 6; ++d) {
        const int cl = nb[d] == XF_FLUID ? 2 : nb[d] == XF_SOLID ? 3 : 0;
        cd |= (uint16_t)(cl << (2 * d));
    }
    double bb = 0.0;
    if (nb[D_XM] != XF_SOLID) bb += face_u(g, u, i, j, k);
    if (nb[D_XP] != XF_SOLID) bb -= face_u(g, u, i + 1, j, k);
    if (nb[D_YM] != XF_SOLID) bb += face_v(g, v, i, j, k);
    if (nb[D_YP] != XF_SOLID) bb -= face_v(g, v, i, j + 1, k);
    if (nb[D_ZM] != XF_SOLID) bb += face_w(g, w, i, j, k);
    if (nb[D_ZP] != XF_SOLID) bb -= face_w(g, w, i, j, k + 1);
    code[c] = cd;
    b[c] = bb;
}
void launch_stencil_rhs(const Geo& g, const uint8_t* mat, const double* u, const double* v, const double* w,
                        uint16_t* code, double* b, hipStream_t st, int op) {
    if (op == OP_NORMAL)
        k_stencil_rhs_normal<<<xf_blocks(g.N), XF_BLOCK, 0, st>>>(g, mat, u, v, w, code, b);
    else
        k_stencil_rhs<<<xf_blocks(g.N), XF_BLOCK, 0, st>>>(g, mat, u, v, w, code, b);
}

// ---------------------------------------------------------------------------
// applyPressure (variational_pressuresolver.cpp:120-204), factor 1.0; cells
// outside the grid count as solid; air ghost pressure = -p(fluid).  Face t of
// the axis's face array.
__device__ inline void apply_face(const Geo& g, const uint8_t* mat, const double* p, double* f, int axis, long t) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1);
    int i, j, k;
    unflat3(t, si, sj, &i, &j, &k);
    int di = axis == 0 ? -1 : 0, dj = axis == 1 ? -1 : 0, dk = axis == 2 ? -1 : 0;
    uint8_t ml = mat_at(g, mat, i + di, j + dj, k + dk, XF_SOLID), mr = mat_at(g, mat, i, j, k, XF_SOLID);
    if (ml != XF_FLUID && mr != XF_FLUID) return;
    double p1, p2;
    if (ml == XF_FLUID && mr == XF_FLUID) {
        p2 = p[cidx(g, i, j, k)];
        p1 = p[cidx(g, i + di, j + dj, k + dk)];
    } else if (ml == XF_FLUID) {
        p1 = p[cidx(g, i + di, j + dj, k + dk)];
        p2 = (mr == XF_SOLID) ? p[cidx(g, i, j, k)] : -p1;
    } else {
        p2 = p[cidx(g, i, j, k)];
        p1 = (ml == XF_SOLID) ? p[cidx(g, i + di, j + dj, k + dk)] : -p2;
    }
    f[t] = f[t] - 1.0 * (p2 - p1);
}
// NormalPressureSolver::applyPressure (pressuresolver.cpp:109-195), factor 1.0:
// outside cells count as solid; an air side has pressure 0; a solid side takes
// the fluid pressure plus (minus) the face velocity, so the update leaves the
// face at u - ((p + u) - p).
__device__ inline void apply_face_normal(const Geo& g, const uint8_t* mat, const double* p, double* f, int axis,
                                         long t) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1);
    int i, j, k;
    unflat3(t, si, sj, &i, &j, &k);
    int di = axis == 0 ? -1 : 0, dj = axis == 1 ? -1 : 0, dk = axis == 2 ? -1 : 0;
    uint8_t ml = mat_at(g, mat, i + di, j + dj, k + dk, XF_SOLID), mr = mat_at(g, mat, i, j, k, XF_SOLID);
    if (ml != XF_FLUID && mr != XF_FLUID) return;
    const double fv = f[t];
    double p1, p2;
    if (ml == XF_FLUID && mr == XF_FLUID) {
        p2 = p[cidx(g, i, j, k)];
        p1 = p[cidx(g, i + di, j + dj, k + dk)];
    } else if (ml == XF_FLUID) {
        p1 = p[cidx(g, i + di, j + dj, k + dk)];
        p2 = (mr == XF_SOLID) ? p1 + fv / 1.0 : 0.0;
    } else {
        p2 = p[cidx(g, i, j, k)];
        p1 = (ml == XF_SOLID) ? p2 - fv / 1.0 : 0.0;
    }
    f[t] = fv - 1.0 * (p2 - p1);
}
__host__ __device__ inline long face_count(const Geo& g, int axis) {
    return (long)(g.ni + (axis == 0)) * (g.nj + (axis == 1)) * (g.nk + (axis == 2));
}
// thread t of a launch over the three face arrays: its axis and its face in that array
__device__ inline bool face3(const Geo& g, long t, int* axis, long* tf) {
    const long n0 = face_count(g, 0), n1 = face_count(g, 1), n2 = face_count(g, 2);
    if (t < n0) { *axis = 0; *tf = t; return true; }
    t -= n0;
    if (t < n1) { *axis = 1; *tf = t; return true; }
    t -= n1;
    *axis = 2;
    *tf = t;
    return t < n2;
}
template <int OP>
__global__ void k_apply_pressure(Geo g, const uint8_t* mat, const double* p, double* fu, double* fv, double* fw) {
    int axis;
    long t;
    if (!face3(g, blockIdx.x * (long)blockDim.x + threadIdx.x, &axis, &t)) return;
    double* f = axis == 0 ? fu : axis == 1 ? fv : fw;
    if (OP == OP_NORMAL) apply_face_normal(g, mat, p, f, axis, t);
    else apply_face(g, mat, p, f, axis, t);
}
void launch_apply_pressure(const Geo& g, const uint8_t* mat, const double* p, double* fu, double* fv, double* fw,
                           hipStream_t st, int op) {
    const long n = face_count(g, 0) + face_count(g, 1) + face_count(g, 2);
    if (op == OP_NORMAL) k_apply_pressure<OP_NORMAL><<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, mat, p, fu, fv, fw);
    else k_apply_pressure<OP_VARIATIONAL><<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, mat, p, fu, fv, fw);
}

// ---------------------------------------------------------------------------
// Velocity extrapolation, FluidSimulator.cpp:511-650.  Layer 0 = faces of
// fluid cells.  The reference grows the layers breadth-first; a face of layer
// L is a still-unmarked face next to a layer L-1 face, and its value is the
// mean over its layer L-1 neighbours (order +i,-i,+j,-j,+k,-k).  Each layer
// reads only finalised layer L-1 values, so one gather kernel per layer gives
// the same values.  In-place marking is safe: an unmarked face only looks for
// neighbours marked L-1, and markers change from -1 to L within the launch.
__device__ inline void extrap_init_face(const Geo& g, const uint8_t* mat, int8_t* mark, int axis, long t) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1);
    int i, j, k;
    unflat3(t, si, sj, &i, &j, &k);
    int oi = axis == 0, oj = axis == 1, ok = axis == 2;
    bool a = mat_at(g, mat, i, j, k, XF_AIR) == XF_FLUID;
    bool b = mat_at(g, mat, i - oi, j - oj, k - ok, XF_AIR) == XF_FLUID;
    mark[t] = (a || b) ? 0 : -1;
}
__global__ void k_extrap_init(Geo g, const uint8_t* mat, int8_t* mark, int axis) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= face_count(g, axis)) return;
    extrap_init_face(g, mat, mark, axis, t);
}
// the three face arrays' markers at mark + axis * mstride
__global__ void k_extrap_init3(Geo g, const uint8_t* mat, int8_t* mark, long mstride) {
    int axis;
    long t;
    if (!face3(g, blockIdx.x * (long)blockDim.x + threadIdx.x, &axis, &t)) return;
    extrap_init_face(g, mat, mark + axis * mstride, axis, t);
}

__device__ inline void extrap_layer_face(const Geo& g, double* f, int8_t* mark, int axis, int L, long t) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1), sk = g.nk + (axis == 2);
    if (mark[t] != -1) return;
    int i, j, k;
    unflat3(t, si, sj, &i, &j, &k);
    const long sij = (long)si * sj;
    double sum = 0.0;
    int count = 0;
    const int8_t want = (int8_t)(L - 1);
    // +i,-i,+j,-j,+k,-k
    if (i + 1 < si && mark[t + 1] == want) { sum += f[t + 1]; count += 1; }
    if (i - 1 >= 0 && mark[t - 1] == want) { sum += f[t - 1]; count += 1; }
    if (j + 1 < sj && mark[t + si] == want) { sum += f[t + si]; count += 1; }
    if (j - 1 >= 0 && mark[t - si] == want) { sum += f[t - si]; count += 1; }
    if (k + 1 < sk && mark[t + sij] == want) { sum += f[t + sij]; count += 1; }
    if (k - 1 >= 0 && mark[t - sij] == want) { sum += f[t - sij]; count += 1; }
    if (count > 0) {
        f[t] = sum / (double)count;
        mark[t] = (int8_t)L;
    }
}
// layer L of the three face arrays in one launch (each array only reads its own
// faces and marker)
__global__ void k_extrap_layer(Geo g, double* fu, double* fv, double* fw, int8_t* mark, long mstride, int L) {
    int axis;
    long t;
    if (!face3(g, blockIdx.x * (long)blockDim.x + threadIdx.x, &axis, &t)) return;
    extrap_layer_face(g, axis == 0 ? fu : axis == 1 ? fv : fw, mark + axis * mstride, axis, L, t);
}

// ---- adjoint of the extrapolation (fixed material) ----
// Layer L holds the faces first reached at step L; each took the mean of its
// neighbours in layer L-1.  The transpose runs the layers backwards: a face of
// layer L-1 gathers g(t) / count(t) from its neighbours t in layer L (count(t)
// = t's neighbours in layer L-1), in the order +i,-i,+j,-j,+k,-k; finally the
// faces of layers >= 1 drop their gradient (their own values were overwritten).
__global__ void k_extrap_mark(Geo g, int8_t* mark, int axis, int L) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1), sk = g.nk + (axis == 2);
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)si * sj * sk || mark[t] != -1) return;
    int i, j, k;
    unflat3(t, si, sj, &i, &j, &k);
    const long sij = (long)si * sj;
    const int8_t want = (int8_t)(L - 1);
    if ((i + 1 < si && mark[t + 1] == want) || (i - 1 >= 0 && mark[t - 1] == want) ||
        (j + 1 < sj && mark[t + si] == want) || (j - 1 >= 0 && mark[t - si] == want) ||
        (k + 1 < sk && mark[t + sij] == want) || (k - 1 >= 0 && mark[t - sij] == want))
        mark[t] = (int8_t)L;
}
__device__ inline int extrap_count(const int8_t* mark, long t, int i, int j, int k, int si, int sj, int sk,
                                   int8_t want) {
    const long sij = (long)si * sj;
    return (i + 1 < si && mark[t + 1] == want) + (i - 1 >= 0 && mark[t - 1] == want) +
           (j + 1 < sj && mark[t + si] == want) + (j - 1 >= 0 && mark[t - si] == want) +
           (k + 1 < sk && mark[t + sij] == want) + (k - 1 >= 0 && mark[t - sij] == want);
}
__global__ void k_extrap_layer_adj(Geo g, double* gf, const int8_t* __restrict__ mark, int axis, int L) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1), sk = g.nk + (axis == 2);
    const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (n >= (long)si * sj * sk || mark[n] != (int8_t)(L - 1)) return;
    int i, j, k;
    unflat3(n, si, sj, &i, &j, &k);
    const long sij = (long)si * sj;
    const int8_t lay = (int8_t)L, prev = (int8_t)(L - 1);
    const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
    double acc = 0.0;
    for (int e = 0; e < 6; ++e) {
        const int a = i + di[e], b = j + dj[e], c = k + dk[e];
        if (a < 0 || b < 0 || c < 0 || a >= si || b >= sj || c >= sk) continue;
        const long t = n + di[e] + (long)si * dj[e] + sij * dk[e];
        if (mark[t] != lay) continue;
        acc += gf[t] / (double)extrap_count(mark, t, a, b, c, si, sj, sk, prev);
    }
    gf[n] += acc;
}
__global__ void k_extrap_drop(long nf, double* gf, const int8_t* __restrict__ mark) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < nf && mark[t] >= 1) gf[t] = 0.0;
}
void launch_extrapolate_adjoint(const Geo& g, const uint8_t* mat, double* gf, int8_t* marker, int axis, int layers,
                                hipStream_t st) {
    const long nf = (long)(g.ni + (axis == 0)) * (g.nj + (axis == 1)) * (g.nk + (axis == 2));
    k_extrap_init<<<xf_blocks(nf), XF_BLOCK, 0, st>>>(g, mat, marker, axis);
    for (int L = 1; L <= layers; ++L) k_extrap_mark<<<xf_blocks(nf), XF_BLOCK, 0, st>>>(g, marker, axis, L);
    for (int L = layers; L >= 1; --L) k_extrap_layer_adj<<<xf_blocks(nf), XF_BLOCK, 0, st>>>(g, gf, marker, axis, L);
    k_extrap_drop<<<xf_blocks(nf), XF_BLOCK, 0, st>>>(nf, gf, marker);
}

void launch_extrapolate(const Geo& g, const uint8_t* mat, double* fu, double* fv, double* fw, int8_t* marker,
                        long mstride, int layers, hipStream_t st) {
    const long n = face_count(g, 0) + face_count(g, 1) + face_count(g, 2);
    k_extrap_init3<<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, mat, marker, mstride);
    for (int L = 1; L <= layers; ++L)
        k_extrap_layer<<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, fu, fv, fw, marker, mstride, L);
}
