// xf_adjoint.hip -- vector-Jacobian products (adjoint kernels) of the substep
// operators.  The reference has no gradients; these kernels are new and are
// checked against an FP64 CPU adjoint, the dot-product identity and central
// finite differences.
//
// Projection (fixed material): u' = u - G A^-1 B u, with B the right-hand-side
// assembly (variational_pressuresolver.cpp:231-326), A the variational matrix
// (symmetric) and G the pressure application (:120-204).  For g = dL/du':
//     dL/du = g + B^T A^-1 (-G^T g)
// -G^T g is a per-cell gather over the cell's six faces (k_apply_adj), the
// transposed solve is a PCG solve with A itself, and B^T is a per-face gather
// over the two adjacent cells (k_rhs_adj).  Both gathers are deterministic.
#include "xf_internal.h"

// pbar = -G^T g: each face f' = f - (p2 - p1), air ghost pressure = -p(fluid),
// cells outside the grid count as solid.
__global__ void k_apply_adj(Geo g, const uint8_t* __restrict__ mat, const double* __restrict__ gu,
                            const double* __restrict__ gv, const double* __restrict__ gw, double* __restrict__ pbar) {
    const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const int i = (int)(c % g.ni), j = (int)((c / g.ni) % g.nj), k = (int)(c / g.nn);
    const uint8_t me = mat[c];
    double acc = 0.0;
    for (int axis = 0; axis < 3; ++axis) {
        const int oi = axis == 0, oj = axis == 1, ok = axis == 2;
        const double* f = axis == 0 ? gu : axis == 1 ? gv : gw;
        const long si = g.ni + oi, sj = g.nj + oj;
        // face on the - side: this cell is the right one
        {
            const uint8_t ml = mat_at(g, mat, i - oi, j - oj, k - ok, XF_SOLID);
            const double x = f[i + si * (j + sj * (long)k)];
            if (ml == XF_FLUID || me == XF_FLUID) {
                if (ml == XF_FLUID && me == XF_AIR) {
                } else if (me == XF_FLUID && ml == XF_AIR) {
                    acc += -2.0 * x;
                } else {
                    acc += -x;
                }
            }
        }
        // face on the + side: this cell is the left one
        {
            const uint8_t mr = mat_at(g, mat, i + oi, j + oj, k + ok, XF_SOLID);
            const double x = f[(i + oi) + si * ((j + oj) + sj * (long)(k + ok))];
            if (me == XF_FLUID || mr == XF_FLUID) {
                if (me == XF_FLUID && mr == XF_AIR) {
                    acc += 2.0 * x;
                } else if (mr == XF_FLUID && me == XF_AIR) {
                } else {
                    acc += x;
                }
            }
        }
    }
    pbar[c] = acc;
}

// g += B^T bbar for one face array: the face is the + face (sign -1) of the
// left cell and the - face (+1) of the right cell; weight 1 towards fluid/air,
// 0.5 towards solid or outside; a solid cell next to fluid has the row
// b[m] += -sign * face * 0.5.
__global__ void k_rhs_adj(Geo g, const uint8_t* __restrict__ mat, const double* __restrict__ bbar, double* f,
                          int axis) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1), sk = g.nk + (axis == 2);
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)si * sj * sk) return;
    const int i = (int)(t % si), j = (int)((t / si) % sj), k = (int)(t / ((long)si * sj));
    const int di = axis == 0, dj = axis == 1, dk = axis == 2;
    const uint8_t OUT = 3;
    const uint8_t ml = mat_at(g, mat, i - di, j - dj, k - dk, OUT), mr = mat_at(g, mat, i, j, k, OUT);
    double acc = 0.0;
    if (ml == XF_FLUID)
        acc += -1.0 * ((mr == XF_FLUID || mr == XF_AIR) ? 1.0 : 0.5) * bbar[cidx(g, i - di, j - dj, k - dk)];
    if (mr == XF_FLUID) acc += ((ml == XF_FLUID || ml == XF_AIR) ? 1.0 : 0.5) * bbar[cidx(g, i, j, k)];
    if (ml == XF_FLUID && mr == XF_SOLID) acc += 0.5 * bbar[cidx(g, i, j, k)];
    if (mr == XF_FLUID && ml == XF_SOLID) acc += -0.5 * bbar[cidx(g, i - di, j - dj, k - dk)];
    f[t] += acc;
}

void launch_apply_adjoint(const Geo& g, const uint8_t* mat, const double* gu, const double* gv, const double* gw,
                          double* pbar, hipStream_t st) {
    k_apply_adj<<<xf_blocks(g.N), XF_BLOCK, 0, st>>>(g, mat, gu, gv, gw, pbar);
}
void launch_rhs_adjoint(const Geo& g, const uint8_t* mat, const double* bbar, double* f, int axis, hipStream_t st) {
    const long nf = (long)(g.ni + (axis == 0)) * (g.nj + (axis == 1)) * (g.nk + (axis == 2));
    k_rhs_adj<<<xf_blocks(nf), XF_BLOCK, 0, st>>>(g, mat, bbar, f, axis);
}
