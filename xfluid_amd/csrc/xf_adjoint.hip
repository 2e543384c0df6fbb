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
#include "../../include/xfluid.h"
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

// ---------------------------------------------------------------------------
// Particle transfers.  Particles are n x 3 arrays (x,y,z interleaved).
//
// G2P sampling v = S(u; p) (_getVelocity, FluidSimulator.cpp:814-894): per
// component a staggered trilinear interpolation.  Its VJP scatters g * corner
// weight to the 8 faces (atomic FP64 adds: the sum order, hence the last bits,
// varies between runs) and adds g * dS/dp to the particle's position gradient.
// Index clamps only move the interpolation base, so dS/dp is the trilinear
// slope times 1/dx; outside the box S = 0 and so are its derivatives.
struct FaceGrads {
    double *u, *v, *w;
};

__device__ inline void trilerp_adj(const double* t, double x, double y, double z, double go, double* gt,
                                   double* gx) {
    const double wx[2] = {1.0 - x, x}, wy[2] = {1.0 - y, y}, wz[2] = {1.0 - z, z};
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = q & 1, b = (q >> 1) & 1, c = q >> 2;
        gt[q] = go * wx[a] * wy[b] * wz[c];
        a0 += t[q] * (a ? 1.0 : -1.0) * wy[b] * wz[c];
        a1 += t[q] * wx[a] * (b ? 1.0 : -1.0) * wz[c];
        a2 += t[q] * wx[a] * wy[b] * (c ? 1.0 : -1.0);
    }
    gx[0] = go * a0; gx[1] = go * a1; gx[2] = go * a2;
}

__device__ void sample_adj(const Geo& g, const Faces& F, const FaceGrads& G, const double* p, const double* go,
                           double* gp) {
    const double dx = g.dx;
    if (!((p[0] >= 0.0) && (p[1] >= 0.0) && (p[2] >= 0.0) && (p[0] <= dx * g.ni) && (p[1] <= dx * g.nj) &&
          (p[2] <= dx * g.nk)))
        return;
    const double invdx = 1.0 / dx;
    for (int comp = 0; comp < 3; ++comp) {
        double x, y, z;
        int i, j, k;
        long si, sj;
        const double* f;
        double* gf;
        if (comp == 0) {
            x = p[0] * invdx; y = (p[1] - 0.5 * dx) * invdx; z = (p[2] - 0.5 * dx) * invdx;
            i = (int)x;
            j = y < 0.0 ? 0 : (int)y; j = j > (g.nj - 2) ? (g.nj - 2) : j;
            k = z < 0.0 ? 0 : (int)z; k = k > (g.nk - 2) ? (g.nk - 2) : k;
            si = g.ni + 1; sj = g.nj; f = F.u; gf = G.u;
        } else if (comp == 1) {
            x = (p[0] - 0.5 * dx) * invdx; y = p[1] * invdx; z = (p[2] - 0.5 * dx) * invdx;
            j = (int)y;
            i = x < 0.0 ? 0 : (int)x; i = i > (g.ni - 2) ? (g.ni - 2) : i;
            k = z < 0.0 ? 0 : (int)z; k = k > (g.nk - 2) ? (g.nk - 2) : k;
            si = g.ni; sj = g.nj + 1; f = F.v; gf = G.v;
        } else {
            z = p[2] * invdx; y = (p[1] - 0.5 * dx) * invdx; x = (p[0] - 0.5 * dx) * invdx;
            k = (int)z;
            j = y < 0.0 ? 0 : (int)y; j = j > (g.nj - 2) ? (g.nj - 2) : j;
            i = x < 0.0 ? 0 : (int)x; i = i > (g.ni - 2) ? (g.ni - 2) : i;
            si = g.ni; sj = g.nj; f = F.w; gf = G.w;
        }
        x -= i; y -= j; z -= k;
        long idx[8];
        double t[8], gt[8], gx[3];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            idx[q] = (i + (q & 1)) + si * ((j + ((q >> 1) & 1)) + sj * (long)(k + (q >> 2)));
            t[q] = f[idx[q]];
        }
        trilerp_adj(t, x, y, z, go[comp], gt, gx);
        if (gf)
            for (int q = 0; q < 8; ++q)
                if (gt[q] != 0.0) atomicAdd(&gf[idx[q]], gt[q]);
        gp[0] += gx[0] * invdx; gp[1] += gx[1] * invdx; gp[2] += gx[2] * invdx;
    }
}

__global__ void k_sample_adj(Geo g, Faces F, FaceGrads G, long n, const double* __restrict__ pos,
                             const double* __restrict__ gvel, double* __restrict__ gpos) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    const double go[3] = {gvel[3 * p], gvel[3 * p + 1], gvel[3 * p + 2]};
    double gp[3] = {0.0, 0.0, 0.0};
    sample_adj(g, F, G, pp, go, gp);
    for (int a = 0; a < 3; ++a) gpos[3 * p + a] += gp[a];
}

// The position update of the advection stage: the RK4 (FluidSimulator.cpp:
// 896-907) / RK2 target np = p + dt * sum b_s k_s, then the collision branch
// (resolve_move, FluidSimulator.cpp:686-722).  The VJP first splits the
// gradient on the new position into the part that reaches the target (grk)
// and the part that goes to p directly (gextra): target accepted: all to the
// target; position kept: all to p; pulled back from the first solid voxel,
// X' = p + rn (t - back) with t = (B - p_a) / rn_a, rn = d / |d|, d = np - p:
//   g_rn = (t - back) g - e_a (g.rn) t / rn_a,  g_d = (I - rn rn^T) g_rn / |d|,
//   grk = g_d,  gextra = g - e_a (g.rn) / rn_a - g_d.
// Then a reverse sweep over the RK stages.  mat == null: no collision branch.
// gpos is overwritten.
__global__ void k_advect_adj(Geo g, Faces F, FaceGrads G, const uint8_t* __restrict__ mat, long n,
                             const double* __restrict__ pos, double dt, int integrator,
                             const double* __restrict__ gnew, double* __restrict__ gpos) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    const double go[3] = {gnew[3 * p], gnew[3 * p + 1], gnew[3 * p + 2]};
    double k1[3], k2[3], k3[3], k4[3], q2[3], q3[3], q4[3], np_[3];
    sample(g, F, pp, k1);
    for (int a = 0; a < 3; ++a) q2[a] = pp[a] + k1[a] * (0.5 * dt);
    sample(g, F, q2, k2);
    if (integrator == XF_INTEGRATOR_RK2) {
        for (int a = 0; a < 3; ++a) np_[a] = pp[a] + k2[a] * dt;
    } else {
        for (int a = 0; a < 3; ++a) q3[a] = pp[a] + k2[a] * (0.5 * dt);
        sample(g, F, q3, k3);
        for (int a = 0; a < 3; ++a) q4[a] = pp[a] + k3[a] * dt;
        sample(g, F, q4, k4);
        for (int a = 0; a < 3; ++a) np_[a] = pp[a] + (k1[a] + k2[a] * 2.0 + k3[a] * 2.0 + k4[a]) * (dt / 6.0);
    }
    double gn[3] = {go[0], go[1], go[2]}, gextra[3] = {0.0, 0.0, 0.0};
    if (mat) {
        double out[3];
        Collision c;
        const int kind = resolve_move(g, mat, pp, np_, out, &c);
        if (kind == MOVE_KEPT) {
            for (int a = 0; a < 3; ++a) { gextra[a] = go[a]; gn[a] = 0.0; }
        } else if (kind == MOVE_PULLED) {
            const int ax = c.axis;
            const double s = go[0] * c.rn[0] + go[1] * c.rn[1] + go[2] * c.rn[2];
            const double cc = c.tmin - c.back;
            double grn[3];
            for (int k = 0; k < 3; ++k) grn[k] = cc * go[k] - (k == ax ? s * c.tmin / c.rn[ax] : 0.0);
            const double rg = c.rn[0] * grn[0] + c.rn[1] * grn[1] + c.rn[2] * grn[2];
            for (int k = 0; k < 3; ++k) {
                gn[k] = (grn[k] - c.rn[k] * rg) / c.l;
                gextra[k] = go[k] - (k == ax ? s / c.rn[ax] : 0.0) - gn[k];
            }
        }
    }
    double gp[3] = {gn[0], gn[1], gn[2]};
    if (integrator == XF_INTEGRATOR_RK2) {
        const double gk2[3] = {gn[0] * dt, gn[1] * dt, gn[2] * dt};
        double gq2[3] = {0.0, 0.0, 0.0}, gk1[3];
        sample_adj(g, F, G, q2, gk2, gq2);
        for (int a = 0; a < 3; ++a) { gp[a] += gq2[a]; gk1[a] = gq2[a] * (0.5 * dt); }
        sample_adj(g, F, G, pp, gk1, gp);
    } else {
        double gk1[3], gk2[3], gk3[3], gk4[3], gq[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a) {
            const double b = gn[a] * (dt / 6.0);
            gk1[a] = b; gk2[a] = b * 2.0; gk3[a] = b * 2.0; gk4[a] = b;
        }
        sample_adj(g, F, G, q4, gk4, gq);
        for (int a = 0; a < 3; ++a) { gp[a] += gq[a]; gk3[a] += gq[a] * dt; gq[a] = 0.0; }
        sample_adj(g, F, G, q3, gk3, gq);
        for (int a = 0; a < 3; ++a) { gp[a] += gq[a]; gk2[a] += gq[a] * (0.5 * dt); gq[a] = 0.0; }
        sample_adj(g, F, G, q2, gk2, gq);
        for (int a = 0; a < 3; ++a) { gp[a] += gq[a]; gk1[a] += gq[a] * (0.5 * dt); }
        sample_adj(g, F, G, pp, gk1, gp);
    }
    for (int a = 0; a < 3; ++a) gpos[3 * p + a] = gp[a] + gextra[a];
}

// P2G (FluidSimulator.cpp:909-1085): u_f = sum_p w_pf v_p / sum_p w_pf (0 when
// the weight sum is <= 1e-7).  VJP for g on the faces, with a_f = g_f / W_f:
//   gvel_p = sum_f a_f w_pf,   gpos_p += sum_f a_f (v_p - u_f) dw_pf/dp.
// Pass 1 scatters the weight and value sums (atomics), pass 2 forms u_f and
// a_f, pass 3 gathers per particle over the faces of its support.
template <class Fn>
__device__ inline void p2g_visit(const Geo& g, int axis, bool tri, const double* pp, Fn&& fn) {
    const double dx = g.dx;
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1), sk = g.nk + (axis == 2);
    const double r = tri ? dx : 1.5 * dx;
    const Wyvill W = make_wyvill(r);
    const double q[3] = {pp[0] + (axis == 0 ? 0.5 * dx : 0.0), pp[1] + (axis == 1 ? 0.5 * dx : 0.0),
                         pp[2] + (axis == 2 ? 0.5 * dx : 0.0)};
    int g0[3], g1[3];
    for (int a = 0; a < 3; ++a) { g0[a] = pos2idx(-r + q[a], dx); g1[a] = pos2idx(q[a] + r, dx); }
    for (int i = g0[0]; i <= g1[0]; ++i)
        for (int j = g0[1]; j <= g1[1]; ++j)
            for (int k = g0[2]; k <= g1[2]; ++k) {
                if (i < 0 || j < 0 || k < 0 || i >= si || j >= sj || k >= sk) continue;
                const double e[3] = {q[0] - idx2pos(i, dx), q[1] - idx2pos(j, dx), q[2] - idx2pos(k, dx)};
                double cw = 0.0, gw[3] = {0.0, 0.0, 0.0};
                if (tri) {
                    const double a[3] = {fabs(e[0]) / dx, fabs(e[1]) / dx, fabs(e[2]) / dx};
                    if (a[0] < 1.0 && a[1] < 1.0 && a[2] < 1.0) {
                        cw = (1.0 - a[0]) * (1.0 - a[1]) * (1.0 - a[2]);
                        for (int t = 0; t < 3; ++t) {
                            const double sg = e[t] > 0 ? 1.0 : e[t] < 0 ? -1.0 : 0.0;
                            gw[t] = -sg / dx * (1.0 - a[(t + 1) % 3]) * (1.0 - a[(t + 2) % 3]);
                        }
                    }
                } else {
                    const double rs = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                    if (rs < W.r2) {
                        cw = 1.0 - W.c0 * rs * rs * rs + W.c1 * rs * rs - W.c2 * rs;
                        const double dw = -3.0 * W.c0 * rs * rs + 2.0 * W.c1 * rs - W.c2;
                        for (int t = 0; t < 3; ++t) gw[t] = dw * 2.0 * e[t];
                    }
                }
                fn(i + (long)si * (j + (long)sj * k), cw, gw);
            }
}

__global__ void k_p2g_sums(Geo g, int axis, bool tri, long n, const double* __restrict__ pos,
                           const double* __restrict__ vel, double* wt, double* val) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    const double vp = vel[3 * p + axis];
    p2g_visit(g, axis, tri, pp, [&](long f, double cw, const double*) {
        if (cw == 0.0) return;
        atomicAdd(&wt[f], cw);
        atomicAdd(&val[f], vp * cw);
    });
}
// in: wt, val (sums), gf (gradient on the faces) -> wt := a_f, val := u_f
__global__ void k_p2g_coef(long nf, const double* __restrict__ gf, double* wt, double* val) {
    const long f = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const double W = wt[f];
    const bool on = W > 0.0000001;
    val[f] = on ? val[f] / W : 0.0;
    wt[f] = on ? gf[f] / W : 0.0;
}
__global__ void k_p2g_adj(Geo g, int axis, bool tri, long n, const double* __restrict__ pos,
                          const double* __restrict__ vel, const double* __restrict__ a, const double* __restrict__ uf,
                          double* __restrict__ gvel, double* __restrict__ gpos) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double pp[3] = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    const double vp = vel[3 * p + axis];
    double gv = 0.0, gp[3] = {0.0, 0.0, 0.0};
    p2g_visit(g, axis, tri, pp, [&](long f, double cw, const double* gw) {
        const double af = a[f];
        if (af == 0.0) return;
        gv += af * cw;
        const double d = af * (vp - uf[f]);
        gp[0] += d * gw[0]; gp[1] += d * gw[1]; gp[2] += d * gw[2];
    });
    gvel[3 * p + axis] = gv;
    for (int t = 0; t < 3; ++t) gpos[3 * p + t] += gp[t];
}

void launch_sample_adjoint(const Geo& g, const double* const F[3], double* const G[3], long n, const double* pos,
                           const double* gvel, double* gpos, hipStream_t st) {
    if (n <= 0) return;
    k_sample_adj<<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, Faces{F[0], F[1], F[2]}, FaceGrads{G[0], G[1], G[2]}, n, pos,
                                                      gvel, gpos);
}
void launch_advect_adjoint(const Geo& g, const double* const F[3], double* const G[3], const uint8_t* mat, long n,
                           const double* pos, double dt, int integrator, const double* gnew, double* gpos,
                           hipStream_t st) {
    if (n <= 0) return;
    k_advect_adj<<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, Faces{F[0], F[1], F[2]}, FaceGrads{G[0], G[1], G[2]}, mat, n,
                                                      pos, dt, integrator, gnew, gpos);
}
// clears the weight and value sums of one face array (a kernel, so a caller
// may record the launch into a graph)
__global__ void k_zero2(long nf, double* __restrict__ a, double* __restrict__ b) {
    const long f = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (f >= nf) return;
    a[f] = 0.0;
    b[f] = 0.0;
}

// scratch: two arrays of the largest face count (their contents on entry do
// not matter: each axis clears what it uses)
void launch_p2g_adjoint(const Geo& g, int kernel, long n, const double* pos, const double* vel,
                        const double* const gf[3], double* gvel, double* gpos, double* s_wt, double* s_val,
                        hipStream_t st) {
    const bool tri = kernel == XF_P2G_TRILINEAR;
    for (int axis = 0; axis < 3; ++axis) {
        const long nf = (long)(g.ni + (axis == 0)) * (g.nj + (axis == 1)) * (g.nk + (axis == 2));
        k_zero2<<<xf_blocks(nf), XF_BLOCK, 0, st>>>(nf, s_wt, s_val);
        if (n > 0) k_p2g_sums<<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, axis, tri, n, pos, vel, s_wt, s_val);
        k_p2g_coef<<<xf_blocks(nf), XF_BLOCK, 0, st>>>(nf, gf[axis], s_wt, s_val);
        if (n > 0) k_p2g_adj<<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, axis, tri, n, pos, vel, s_wt, s_val, gvel, gpos);
    }
}

// ---------------------------------------------------------------------------
// VJP of the particle-blob density (DirectedDistanceField::updateFeild,
// directeddistancefield.cpp:122-169, 300-328): phi(v) = sum_p W(|x_v - p|^2),
// W(s) = 1 - c0 s^3 + c1 s^2 - c2 s for s < R^2, x_v = (v + 1/2) dx/d.
// gpos_p += sum_v gphi(v) W'(s) (-2)(x_v - p), W'(s) = -3 c0 s^2 + 2 c1 s - c2,
// over the vertex box of the forward scatter (pos2idx(p -+ R) scaled by d), in
// its i, j, k loop order.  One thread per particle.
__global__ void k_density_adj(Geo g, int d, double R, long n, const double* __restrict__ pos,
                              const double* __restrict__ gphi, double* __restrict__ gpos) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int si = d * g.ni + 1, sj = d * g.nj + 1, sk = d * g.nk + 1;
    const double R2 = R * R;
    const double c0 = (4.0 / 9.0) * (1.0 / (R2 * R2 * R2)), c1 = (17.0 / 9.0) * (1.0 / (R2 * R2)),
                 c2 = (22.0 / 9.0) * (1.0 / R2);
    const double subdx = g.dx / (double)d;
    const double x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
    const int g0i = pos2idx(-R + x, g.dx) * d, g0j = pos2idx(-R + y, g.dx) * d, g0k = pos2idx(-R + z, g.dx) * d;
    const int g1i = (pos2idx(x + R, g.dx) + 1) * d, g1j = (pos2idx(y + R, g.dx) + 1) * d,
              g1k = (pos2idx(z + R, g.dx) + 1) * d;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int i = g0i; i <= g1i; ++i)
        for (int j = g0j; j <= g1j; ++j)
            for (int k = g0k; k <= g1k; ++k) {
                if (i < 0 || j < 0 || k < 0 || i >= si || j >= sj || k >= sk) continue;
                const double ex = idx2pos(i, subdx) - x, ey = idx2pos(j, subdx) - y, ez = idx2pos(k, subdx) - z;
                const double rs = ex * ex + ey * ey + ez * ez;
                if (!(rs < R2)) continue;
                const double dw = -3.0 * c0 * rs * rs + 2.0 * c1 * rs - c2;
                const double f = gphi[i + (long)si * (j + (long)sj * k)] * dw * -2.0;
                ax += f * ex;
                ay += f * ey;
                az += f * ez;
            }
    gpos[3 * p] += ax;
    gpos[3 * p + 1] += ay;
    gpos[3 * p + 2] += az;
}

void launch_density_adjoint(const Geo& g, int d, double R, long n, const double* pos, const double* gphi,
                            double* gpos, hipStream_t st) {
    if (n <= 0) return;
    k_density_adj<<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, d, R, n, pos, gphi, gpos);
}
