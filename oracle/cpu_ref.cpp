// cpu_ref.cpp -- FP64 CPU restatement of xFluid's substep (the oracle).
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, and only as the checker
// (or the timed CPU baseline), never as the product path.
//
// Every function restates the reference's algorithm with the same arithmetic
// order (compile with -ffp-contract=off, no FMA), so that on identical inputs
// it reproduces the reference bit for bit.  Citations are to files under
// xFluid/xFluid/ of the reference (variationalpressure/ for the solver files).
//
// Parity status of each part (how it is pinned) is listed in tests/ and in the
// header comment of each section below.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <climits>
#include <string>
#include "../include/xfluid.h"
#include "../xfluid_amd/csrc/xf_mc_tables.h"

namespace {

enum : uint8_t { AIR = 0, FLUID = 1, SOLID = 2 };

// ---------------------------------------------------------------------------
// Index helpers: xfluidutil.cpp:3-21 (pos2Index truncates, idx2Pos = (i+.5)dx)
inline int pos2idx(double p, double dx) { return (int)(p / dx); }
inline double idx2pos(int i, double dx) { return (i + 0.5) * dx; }

// xfluidutil.cpp:53-65
inline double trilerp(const double* p, double x, double y, double z) {
    double v0 = (1.0 - x) * p[0] + x * p[1];
    double v1 = (1.0 - x) * p[2] + x * p[3];
    double v2 = (1.0 - x) * p[4] + x * p[5];
    double v3 = (1.0 - x) * p[6] + x * p[7];
    v0 = (1.0 - y) * v0 + y * v1;
    v1 = (1.0 - y) * v2 + y * v3;
    return (1.0 - z) * v0 + z * v1;
}

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double u01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// ---------------------------------------------------------------------------
// Variational pressure system on the 7-point stencil.
// A is held as 7 coefficient planes, entry order = the column order of a CSR
// row (sparse_matrix.h:266-276): [c-nn, c-ni, c-1, c, c+1, c+ni, c+nn].
enum { A_ZM = 0, A_YM = 1, A_XM = 2, A_D = 3, A_XP = 4, A_YP = 5, A_ZP = 6 };

// variational_pressuresolver.cpp:231-326.  Fluid cells visited with i slowest,
// k fastest; neighbour order +i,-i,+j,-j,+k,-k (xfluidutil.cpp:33-41).
void assemble(int ni, int nj, int nk, const uint8_t* mat, const double* u, const double* v,
              const double* w, double* A, double* b) {
    const long N = (long)ni * nj * nk, nn = (long)ni * nj;
    std::fill(A, A + 7 * N, 0.0);
    std::fill(b, b + N, 0.0);
    auto U = [&](int i, int j, int k) { return u[i + (long)(ni + 1) * (j + (long)nj * k)]; };
    auto V = [&](int i, int j, int k) { return v[i + (long)ni * (j + (long)(nj + 1) * k)]; };
    auto W = [&](int i, int j, int k) { return w[i + (long)ni * (j + (long)nj * k)]; };
    const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
    // slot of neighbour m in row c, and of c in row m
    const int slot_cm[6] = {A_XP, A_XM, A_YP, A_YM, A_ZP, A_ZM};
    const int slot_mc[6] = {A_XM, A_XP, A_YM, A_YP, A_ZM, A_ZP};
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k) {
                long c = i + (long)ni * j + nn * k;
                if (mat[c] != FLUID) continue;
                for (int e = 0; e < 6; ++e) {
                    int gi = i + di[e], gj = j + dj[e], gk = k + dk[e];
                    double vsign = 1.0, face;
                    if (e < 2) { int fi = i + (e == 0 ? 1 : 0); if (e == 0) vsign = -1.0; face = U(fi, j, k); }
                    else if (e < 4) { int fj = j + (e == 2 ? 1 : 0); if (e == 2) vsign = -1.0; face = V(i, fj, k); }
                    else { int fk = k + (e == 4 ? 1 : 0); if (e == 4) vsign = -1.0; face = W(i, j, fk); }
                    long m = gi + (long)ni * gj + nn * gk;
                    uint8_t mm = mat[m];
                    if (mm == FLUID) {
                        A[slot_cm[e] * N + c] += -1.0;
                        A[A_D * N + c] += 1.0;
                        b[c] += vsign * face;
                    } else if (mm == AIR) {
                        A[A_D * N + c] += 2.0;
                        b[c] += vsign * face;
                    } else {
                        A[slot_cm[e] * N + c] += -0.5;
                        A[A_D * N + c] += 0.5;
                        b[c] += vsign * face * 0.5;
                        A[slot_mc[e] * N + m] += -0.5;
                        A[A_D * N + m] += 0.5;
                        b[m] += -vsign * face * 0.5;
                    }
                }
            }
}

// NormalPressureSolver::_calculate_coefficient (pressuresolver.cpp:220-347):
// rows for fluid cells only (i slowest), neighbour order +i,-i,+j,-j,+k,-k
// (xfluidutil.cpp:33-41); an in-range non-solid neighbour adds 1 to the
// diagonal, a fluid one also -1 off the diagonal; then b, per fluid cell,
// +u(i) -u(i+1) +v(j) -v(j+1) +w(k) -w(k+1) for the sides whose cell is in
// range and not solid.
void assemble_normal(int ni, int nj, int nk, const uint8_t* mat, const double* u, const double* v,
                     const double* w, double* A, double* b) {
    const long N = (long)ni * nj * nk, nn = (long)ni * nj;
    std::fill(A, A + 7 * N, 0.0);
    std::fill(b, b + N, 0.0);
    auto in = [&](int i, int j, int k) { return i >= 0 && j >= 0 && k >= 0 && i < ni && j < nj && k < nk; };
    auto Mt = [&](int i, int j, int k) -> uint8_t { return in(i, j, k) ? mat[i + (long)ni * j + nn * k] : SOLID; };
    auto U = [&](int i, int j, int k) { return u[i + (long)(ni + 1) * (j + (long)nj * k)]; };
    auto V = [&](int i, int j, int k) { return v[i + (long)ni * (j + (long)(nj + 1) * k)]; };
    auto W = [&](int i, int j, int k) { return w[i + (long)ni * (j + (long)nj * k)]; };
    const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
    const int slot_cm[6] = {A_XP, A_XM, A_YP, A_YM, A_ZP, A_ZM};
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k) {
                long c = i + (long)ni * j + nn * k;
                if (mat[c] != FLUID) continue;
                int count = 0;
                for (int e = 0; e < 6; ++e) {
                    int gi = i + di[e], gj = j + dj[e], gk = k + dk[e];
                    if (!in(gi, gj, gk)) continue;
                    uint8_t mm = Mt(gi, gj, gk);
                    if (mm != SOLID) {
                        ++count;
                        if (mm == FLUID) A[slot_cm[e] * N + c] += -1.0;
                    }
                }
                A[A_D * N + c] += (double)count * 1.0;
            }
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nk; ++k) {
                long c = i + (long)ni * j + nn * k;
                if (mat[c] != FLUID) continue;
                if (Mt(i - 1, j, k) != SOLID) b[c] += U(i, j, k);
                if (Mt(i + 1, j, k) != SOLID) b[c] -= U(i + 1, j, k);
                if (Mt(i, j - 1, k) != SOLID) b[c] += V(i, j, k);
                if (Mt(i, j + 1, k) != SOLID) b[c] -= V(i, j + 1, k);
                if (Mt(i, j, k - 1) != SOLID) b[c] += W(i, j, k);
                if (Mt(i, j, k + 1) != SOLID) b[c] -= W(i, j, k + 1);
            }
}

// sparse_matrix.h:266-276 (multiply): present entries in column order.
void spmv(int ni, int nj, int nk, const double* A, const double* x, double* y) {
    const long N = (long)ni * nj * nk, nn = (long)ni * nj;
    const long off[7] = {-nn, -(long)ni, -1, 0, 1, (long)ni, nn};
    for (long c = 0; c < N; ++c) {
        double s = 0;
        for (int e = 0; e < 7; ++e) {
            double a = A[e * N + c];
            if (a != 0.0) s += a * x[c + off[e]];
        }
        y[c] = s;
    }
}

// pcg_solver.h:85-186 specialised to the 7-point pattern (all dims >= 3): the
// other entries of column k are never in row j's pattern, so each update of
// pivot j from column k is  d_j -= L*L;  d_j -= (tau*L)*missing.
void mic0(int ni, int nj, int nk, const double* A, double tau, double sigma, double* inv) {
    const long N = (long)ni * nj * nk, nn = (long)ni * nj;
    std::vector<double> d(A + A_D * N, A + A_D * N + N);
    std::fill(inv, inv + N, 0.0);
    for (long k = 0; k < N; ++k) {
        double akk = A[A_D * N + k];
        if (akk == 0) continue;
        double ik = (d[k] < sigma * akk) ? 1 / std::sqrt(akk) : 1 / std::sqrt(d[k]);
        inv[k] = ik;
        // lower entries of column k = upper entries of row k (symmetric), ascending
        long rows[3];
        double L[3];
        int cnt = 0;
        const int slots[3] = {A_XP, A_YP, A_ZP};
        const long offs[3] = {1, (long)ni, nn};
        for (int e = 0; e < 3; ++e) {
            double a = A[slots[e] * N + k];
            if (a != 0.0) { rows[cnt] = k + offs[e]; L[cnt] = a * ik; ++cnt; }
        }
        for (int p = 0; p < cnt; ++p) {
            long j = rows[p];
            double mult = L[p];
            double missing = 0;
            for (int a = 0; a < cnt; ++a) if (rows[a] < j) missing += L[a];
            d[j] -= mult * L[p];
            for (int a = 0; a < cnt; ++a) if (rows[a] > j) missing += L[a];
            d[j] -= tau * mult * missing;
        }
    }
}

// pcg_solver.h:192-204 (solve_lower), column-oriented as in the reference.
void sweep_fwd(int ni, int nj, int nk, const double* A, const double* inv, const double* r, double* q) {
    const long N = (long)ni * nj * nk, nn = (long)ni * nj;
    std::copy(r, r + N, q);
    const int slots[3] = {A_XP, A_YP, A_ZP};
    const long offs[3] = {1, (long)ni, nn};
    for (long i = 0; i < N; ++i) {
        q[i] *= inv[i];
        for (int e = 0; e < 3; ++e) {
            double a = A[slots[e] * N + i];
            if (a != 0.0) q[i + offs[e]] -= (a * inv[i]) * q[i];
        }
    }
}

// pcg_solver.h:207-220 (solve_lower_transpose_in_place)
void sweep_bwd(int ni, int nj, int nk, const double* A, const double* inv, double* x) {
    const long N = (long)ni * nj * nk, nn = (long)ni * nj;
    const int slots[3] = {A_XP, A_YP, A_ZP};
    const long offs[3] = {1, (long)ni, nn};
    long i = N;
    do {
        --i;
        for (int e = 0; e < 3; ++e) {
            double a = A[slots[e] * N + i];
            if (a != 0.0) x[i] -= (a * inv[i]) * x[i + offs[e]];
        }
        x[i] *= inv[i];
    } while (i != 0);
}

double dot(const double* x, const double* y, long n) {
    double s = 0;
    for (long i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
}
double absmax(const double* x, long n) {
    double m = 0;
    for (long i = 0; i < n; ++i) if (std::fabs(x[i]) > m) m = std::fabs(x[i]);
    return m;
}

// pcg_solver.h:243-285.  Returns 0 converged, 1 not converged, 2 rho invalid.
int pcg(int ni, int nj, int nk, const double* A, const double* b, double* x, double tol_factor,
        int max_iter, double tau, double sigma, int* iters_out, double* res_out, double* inv_out) {
    const long N = (long)ni * nj * nk;
    std::vector<double> r(b, b + N), z(N), s(N), inv(N);
    std::fill(x, x + N, 0.0);
    if (tol_factor < 1e-30) tol_factor = 1e-30;
    double res = absmax(r.data(), N);
    *res_out = res;
    if (res == 0) { *iters_out = 0; if (inv_out) std::fill(inv_out, inv_out + N, 0.0); return 0; }
    double tol = tol_factor * res;
    mic0(ni, nj, nk, A, tau, sigma, inv.data());
    if (inv_out) std::copy(inv.begin(), inv.end(), inv_out);
    sweep_fwd(ni, nj, nk, A, inv.data(), r.data(), z.data());
    sweep_bwd(ni, nj, nk, A, inv.data(), z.data());
    double rho = dot(z.data(), r.data(), N);
    if (rho == 0 || rho != rho) { *iters_out = 0; return 2; }
    s = z;
    int it;
    for (it = 0; it < max_iter; ++it) {
        spmv(ni, nj, nk, A, s.data(), z.data());
        double alpha = rho / dot(s.data(), z.data(), N);
        for (long i = 0; i < N; ++i) x[i] += alpha * s[i];
        for (long i = 0; i < N; ++i) r[i] += -alpha * z[i];
        res = absmax(r.data(), N);
        *res_out = res;
        if (res <= tol) { *iters_out = it + 1; return 0; }
        sweep_fwd(ni, nj, nk, A, inv.data(), r.data(), z.data());
        sweep_bwd(ni, nj, nk, A, inv.data(), z.data());
        double rho_new = dot(z.data(), r.data(), N);
        double beta = rho_new / rho;
        for (long i = 0; i < N; ++i) z[i] += beta * s[i];
        s.swap(z);
        rho = rho_new;
    }
    *iters_out = it;
    return 1;
}

// variational_pressuresolver.cpp:120-204 (factor = 1.0)
void apply_pressure(int ni, int nj, int nk, const uint8_t* mat, const double* p, double* u, double* v,
                    double* w) {
    const long nn = (long)ni * nj;
    auto M = [&](int i, int j, int k) -> uint8_t {
        if (i < 0 || j < 0 || k < 0 || i >= ni || j >= nj || k >= nk) return SOLID;
        return mat[i + (long)ni * j + nn * k];
    };
    auto P = [&](int i, int j, int k) { return p[i + (long)ni * j + nn * k]; };
    for (int axis = 0; axis < 3; ++axis) {
        int di = axis == 0 ? -1 : 0, dj = axis == 1 ? -1 : 0, dk = axis == 2 ? -1 : 0;
        int si = ni + (axis == 0), sj = nj + (axis == 1), sk = nk + (axis == 2);
        double* f = axis == 0 ? u : axis == 1 ? v : w;
#pragma omp parallel for
        for (int i = 0; i < si; ++i)
            for (int j = 0; j < sj; ++j)
                for (int k = 0; k < sk; ++k) {
                    uint8_t ml = M(i + di, j + dj, k + dk), mr = M(i, j, k);
                    if (ml != FLUID && mr != FLUID) continue;
                    double p1, p2;
                    if (ml == FLUID && mr == FLUID) { p2 = P(i, j, k); p1 = P(i + di, j + dj, k + dk); }
                    else if (ml == FLUID) {
                        p1 = P(i + di, j + dj, k + dk);
                        p2 = (mr == SOLID) ? P(i, j, k) : -p1;
                    } else {
                        p2 = P(i, j, k);
                        p1 = (ml == SOLID) ? P(i + di, j + dj, k + dk) : -p2;
                    }
                    long fi = i + (long)si * (j + (long)sj * k);
                    f[fi] = f[fi] - 1.0 * (p2 - p1);
                }
    }
}

// NormalPressureSolver::applyPressure (pressuresolver.cpp:109-195), factor
// 1.0: outside cells count as solid; an air side has pressure 0, a solid side
// the fluid pressure plus (minus) the face velocity.
void apply_pressure_normal(int ni, int nj, int nk, const uint8_t* mat, const double* p, double* u, double* v,
                           double* w) {
    const long nn = (long)ni * nj;
    auto M = [&](int i, int j, int k) -> uint8_t {
        if (i < 0 || j < 0 || k < 0 || i >= ni || j >= nj || k >= nk) return SOLID;
        return mat[i + (long)ni * j + nn * k];
    };
    auto P = [&](int i, int j, int k) { return p[i + (long)ni * j + nn * k]; };
    const double factor = 1.0;
    for (int axis = 0; axis < 3; ++axis) {
        int di = axis == 0 ? -1 : 0, dj = axis == 1 ? -1 : 0, dk = axis == 2 ? -1 : 0;
        int si = ni + (axis == 0), sj = nj + (axis == 1), sk = nk + (axis == 2);
        double* f = axis == 0 ? u : axis == 1 ? v : w;
        for (int i = 0; i < si; ++i)
            for (int j = 0; j < sj; ++j)
                for (int k = 0; k < sk; ++k) {
                    uint8_t ml = M(i + di, j + dj, k + dk), mr = M(i, j, k);
                    if (ml != FLUID && mr != FLUID) continue;
                    long fi = i + (long)si * (j + (long)sj * k);
                    double p1 = 0.0, p2 = 0.0;
                    if (ml == FLUID && mr == FLUID) { p2 = P(i, j, k); p1 = P(i + di, j + dj, k + dk); }
                    else if (ml == FLUID) {
                        p1 = P(i + di, j + dj, k + dk);
                        p2 = (mr == SOLID) ? p1 + f[fi] / factor : 0.0;
                    } else {
                        p2 = P(i, j, k);
                        p1 = (ml == SOLID) ? p2 - f[fi] / factor : 0.0;
                    }
                    f[fi] = f[fi] - factor * (p2 - p1);
                }
    }
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Adjoint of the projection u' = u - G A^-1 B u for fixed material (no
// reference counterpart: the reference has no gradients).  B is the RHS
// assembly above (b = B u), G the pressure application; A is symmetric, so
// the transposed solve is a solve with A.  Given g = dL/du', returns
// dL/du = g + B^T A^-1 (-G^T g), computed in FP64 with the pcg() above.

// pbar = -G^T g (transpose of apply_pressure; outside the grid counts as solid)
void apply_pressure_adjoint(int ni, int nj, int nk, const uint8_t* mat, const double* gu, const double* gv,
                            const double* gw, double* pbar) {
    const long N = (long)ni * nj * nk, nn = (long)ni * nj;
    std::fill(pbar, pbar + N, 0.0);
    auto M = [&](int i, int j, int k) -> uint8_t {
        if (i < 0 || j < 0 || k < 0 || i >= ni || j >= nj || k >= nk) return SOLID;
        return mat[i + (long)ni * j + nn * k];
    };
    auto add = [&](int i, int j, int k, double x) {
        if (i < 0 || j < 0 || k < 0 || i >= ni || j >= nj || k >= nk) return;
        pbar[i + (long)ni * j + nn * k] += x;
    };
    for (int axis = 0; axis < 3; ++axis) {
        int di = axis == 0 ? -1 : 0, dj = axis == 1 ? -1 : 0, dk = axis == 2 ? -1 : 0;
        int si = ni + (axis == 0), sj = nj + (axis == 1), sk = nk + (axis == 2);
        const double* g = axis == 0 ? gu : axis == 1 ? gv : gw;
        for (int k = 0; k < sk; ++k)
            for (int j = 0; j < sj; ++j)
                for (int i = 0; i < si; ++i) {
                    uint8_t ml = M(i + di, j + dj, k + dk), mr = M(i, j, k);
                    if (ml != FLUID && mr != FLUID) continue;
                    const double x = g[i + (long)si * (j + (long)sj * k)];
                    // f' = f - (p2 - p1); air ghost p = -p(fluid)
                    if (ml == FLUID && mr == AIR) add(i + di, j + dj, k + dk, 2.0 * x);
                    else if (mr == FLUID && ml == AIR) add(i, j, k, -2.0 * x);
                    else { add(i, j, k, -x); add(i + di, j + dj, k + dk, x); }
                }
    }
}

// g += B^T bbar (transpose of the RHS part of assemble())
void rhs_adjoint(int ni, int nj, int nk, const uint8_t* mat, const double* bbar, double* gu, double* gv,
                 double* gw) {
    const long nn = (long)ni * nj;
    const uint8_t OUT = 3;
    auto M = [&](int i, int j, int k) -> uint8_t {
        if (i < 0 || j < 0 || k < 0 || i >= ni || j >= nj || k >= nk) return OUT;
        return mat[i + (long)ni * j + nn * k];
    };
    auto B = [&](int i, int j, int k) { return bbar[i + (long)ni * j + nn * k]; };
    for (int axis = 0; axis < 3; ++axis) {
        int di = axis == 0 ? -1 : 0, dj = axis == 1 ? -1 : 0, dk = axis == 2 ? -1 : 0;
        int si = ni + (axis == 0), sj = nj + (axis == 1), sk = nk + (axis == 2);
        double* g = axis == 0 ? gu : axis == 1 ? gv : gw;
        for (int k = 0; k < sk; ++k)
            for (int j = 0; j < sj; ++j)
                for (int i = 0; i < si; ++i) {
                    const uint8_t ml = M(i + di, j + dj, k + dk), mr = M(i, j, k);
                    double acc = 0.0;
                    // the face is the + face (vsign -1) of the left cell, the - face (+1) of the right
                    if (ml == FLUID) acc += -1.0 * ((mr == FLUID || mr == AIR) ? 1.0 : 0.5) * B(i + di, j + dj, k + dk);
                    if (mr == FLUID) acc += ((ml == FLUID || ml == AIR) ? 1.0 : 0.5) * B(i, j, k);
                    // rows of fluid-adjacent solid cells: b[m] += -vsign * face * 0.5
                    if (ml == FLUID && mr == SOLID) acc += 0.5 * B(i, j, k);
                    if (mr == FLUID && ml == SOLID) acc += -0.5 * B(i + di, j + dj, k + dk);
                    g[i + (long)si * (j + (long)sj * k)] += acc;
                }
    }
}

struct Sim {
    int ni, nj, nk;
    double dx = 0.0001;
    xf_params prm;
    std::vector<double> u, v, w, us, vs, ws;
    std::vector<uint8_t> mat;
    std::vector<double> px, py, pz, vx, vy, vz;
    std::vector<double> phi;
    std::vector<double> pressure, rhs, micinv;
    std::vector<double> forces;  // triples
    std::vector<int32_t> surf;   // sub-surface cells (flat sub-cell index)
    std::vector<uint8_t> cube;   // their marching-cubes case
    xf_solve_stats last{};
    int frame = 0;
    // fluid sources (fluidsource.h:10-53): cells fixed at creation, velocity
    struct Source {
        std::vector<long> cells;
        double v[3];
        double normv;
    };
    std::vector<Source> sources;
    uint64_t src_state = 0;  // 31-bit SplitMix64 stream in place of rand()

    long N() const { return (long)ni * nj * nk; }
    long nu() const { return (long)(ni + 1) * nj * nk; }
    long nv() const { return (long)ni * (nj + 1) * nk; }
    long nw() const { return (long)ni * nj * (nk + 1); }
    int d() const { return prm.sub_division; }
    long nphi() const { return (long)(d() * ni + 1) * (d() * nj + 1) * (d() * nk + 1); }

    Sim(int a, int b, int c) : ni(a), nj(b), nk(c) {
        xf_params_default(&prm);
        u.assign(nu(), 0.0); v.assign(nv(), 0.0); w.assign(nw(), 0.0);
        us = u; vs = v; ws = w;
        mat.assign(N(), AIR);
        pressure.assign(N(), 0.0); rhs.assign(N(), 0.0); micinv.assign(N(), 0.0);
        phi.assign(nphi(), 0.0);
    }

    uint8_t& M(int i, int j, int k) { return mat[i + (long)ni * j + (long)ni * nj * k]; }
    bool mat_in(int i, int j, int k) const { return i >= 0 && j >= 0 && k >= 0 && i < ni && j < nj && k < nk; }

    // FluidSimulator.cpp:380-385
    double particle_radius() const {
        double volume = dx * dx * dx / 8.0;
        double pi = 3.141592653;
        return std::pow(3 * volume / (4 * pi), 1.0 / 3.0);
    }
    double blob_radius() const { return prm.blob_radius > 0 ? prm.blob_radius : particle_radius() * 3.0; }

    // FluidSimulator.cpp:64-96
    void set_solid_boundary() {
        for (int i = 0; i < ni; i += (ni - 1)) for (int j = 0; j < nj; ++j) for (int k = 0; k < nk; ++k) M(i, j, k) = SOLID;
        for (int i = 0; i < ni; ++i) for (int j = 0; j < nj; j += (nj - 1)) for (int k = 0; k < nk; ++k) M(i, j, k) = SOLID;
        for (int i = 0; i < ni; ++i) for (int j = 0; j < nj; ++j) for (int k = 0; k < nk; k += (nk - 1)) M(i, j, k) = SOLID;
    }
    // FluidSimulator.cpp:103-121 (note the float r2)
    void add_sphere(double x, double y, double z, double r) {
        float r2 = r * r;
        for (int i = 1; i < ni - 1; ++i) for (int j = 1; j < nj - 1; ++j) for (int k = 1; k < nk - 1; ++k) {
            double gx = idx2pos(i, dx), gy = idx2pos(j, dx), gz = idx2pos(k, dx);
            double ex = x - gx, ey = y - gy, ez = z - gz;
            if (ex * ex + ey * ey + ez * ez <= r2) M(i, j, k) = FLUID;
        }
    }
    // FluidSimulator.cpp:123-140
    void add_cubic(const double* lu, const double* rd) {
        for (int i = 1; i < ni - 1; ++i) for (int j = 1; j < nj - 1; ++j) for (int k = 1; k < nk - 1; ++k) {
            double gx = idx2pos(i, dx), gy = idx2pos(j, dx), gz = idx2pos(k, dx);
            if ((gx <= rd[0] && gx >= lu[0]) && (gy <= rd[1] && gy >= lu[1]) && (gz <= rd[2] && gz >= lu[2]))
                M(i, j, k) = FLUID;
        }
    }
    // FluidSimulator.cpp:168-184, 347-370 with a seeded generator, or (draws
    // given, reference mode) with the jitter randomDouble(-0.01, 0.01) of the
    // given 31-bit values, three per particle in x, y, z order
    void seed(int mode, uint64_t s, const uint32_t* draws = nullptr) {
        long d = 0;
        px.clear(); py.clear(); pz.clear(); vx.clear(); vy.clear(); vz.clear();
        const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1},
                              {-1, 1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, 1}};
        double q = 0.25 * dx;
        for (int i = 0; i < ni; ++i) for (int j = 0; j < nj; ++j) for (int k = 0; k < nk; ++k) {
            if (M(i, j, k) != FLUID) continue;
            double c[3] = {idx2pos(i, dx), idx2pos(j, dx), idx2pos(k, dx)};
            for (int t = 0; t < 8; ++t) {
                double p[3];
                for (int a = 0; a < 3; ++a) {
                    if (draws) {
                        double pt = sg[t][a] < 0 ? c[a] - q : c[a] + q;
                        p[a] = pt + random_double(-0.01, 0.01, draws[d++]);
                        continue;
                    }
                    double r = u01(s);
                    if (mode == XF_SEED_REFERENCE) {
                        double pt = sg[t][a] < 0 ? c[a] - q : c[a] + q;
                        p[a] = pt + (-0.01 + r * (0.01 - -0.01));
                    } else {
                        double lo = sg[t][a] < 0 ? c[a] - 0.5 * dx : c[a];
                        p[a] = lo + r * (0.5 * dx);
                    }
                }
                px.push_back(p[0]); py.push_back(p[1]); pz.push_back(p[2]);
                vx.push_back(0.0); vy.push_back(0.0); vz.push_back(0.0);
            }
        }
    }

    // ---- a3: directeddistancefield.cpp:27-37, 122-169, 246-256, 300-328 ----
    void density() {
        const int dd = d();
        const int si = dd * ni + 1, sj = dd * nj + 1, sk = dd * nk + 1;
        const long NP = (long)si * sj * sk;
        const double R = blob_radius(), R2 = R * R;
        const double c0 = (4.0 / 9.0) * (1.0 / (R2 * R2 * R2));
        const double c1 = (17.0 / 9.0) * (1.0 / (R2 * R2));
        const double c2 = (22.0 / 9.0) * (1.0 / (R2));
        const double subdx = dx / (double)dd;
        const int T = 6;
        const long n = (long)px.size();
        const long npt = n / T + 1;
        std::vector<std::vector<double>> phis(T);
#pragma omp parallel for
        for (int t = 0; t < T; ++t) {
            phis[t].assign(NP, 0.0);
            double* ph = phis[t].data();
            long s = t * npt, e = std::min(s + npt, n);
            for (long p = s; p < e; ++p) {
                double x = px[p], y = py[p], z = pz[p];
                int g0i = pos2idx(-R + x, dx) * dd, g0j = pos2idx(-R + y, dx) * dd, g0k = pos2idx(-R + z, dx) * dd;
                int g1i = (pos2idx(x + R, dx) + 1) * dd, g1j = (pos2idx(y + R, dx) + 1) * dd,
                    g1k = (pos2idx(z + R, dx) + 1) * dd;
                for (int i = g0i; i <= g1i; ++i)
                    for (int j = g0j; j <= g1j; ++j)
                        for (int k = g0k; k <= g1k; ++k) {
                            if (i < 0 || j < 0 || k < 0 || i >= si || j >= sj || k >= sk) continue;
                            double ex = idx2pos(i, subdx) - x, ey = idx2pos(j, subdx) - y, ez = idx2pos(k, subdx) - z;
                            double rs = ex * ex + ey * ey + ez * ez;
                            double wv = rs < R2 ? 1.0 - c0 * rs * rs * rs + c1 * rs * rs - c2 * rs : 0.0;
                            ph[i + (long)si * (j + (long)sj * k)] += wv;
                        }
            }
        }
#pragma omp parallel for
        for (long c = 0; c < NP; ++c) {
            double s = 0.0;
            for (int t = 0; t < T; ++t) s += phis[t][c];
            phi[c] = s;
        }
    }

    // directeddistancefield.cpp:171-244, 330-352 (surface list);
    // fluidmesher.cpp:397-417 (cube index, strict >)
    void surface() {
        const int dd = d();
        const int si = dd * ni + 1, sj = dd * nj + 1;
        const double thr = prm.iso;
        auto PHI = [&](int i, int j, int k) { return phi[i + (long)si * (j + (long)sj * k)]; };
        auto is_surface = [&](int i, int j, int k, int span) {
            bool in = false, out = false;
            for (int a = i; a <= i + span; ++a) for (int b = j; b <= j + span; ++b) for (int c = k; c <= k + span; ++c) {
                if (PHI(a, b, c) >= thr) in = true; else out = true;
            }
            return in && out;
        };
        surf.clear(); cube.clear();
        const int sni = dd * ni, snj = dd * nj;
        for (int i = 0; i < ni; ++i) for (int j = 0; j < nj; ++j) for (int k = 0; k < nk; ++k) {
            int s0 = i * dd, s1 = j * dd, s2 = k * dd;
            if (!is_surface(s0, s1, s2, dd)) continue;
            for (int a = s0; a < s0 + dd; ++a) for (int b = s1; b < s1 + dd; ++b) for (int c = s2; c < s2 + dd; ++c) {
                if (!is_surface(a, b, c, 1)) continue;
                surf.push_back((int32_t)(a + (long)sni * (b + (long)snj * c)));
                const int cv[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
                int ci = 0;
                for (int t = 0; t < 8; ++t) if (PHI(a + cv[t][0], b + cv[t][1], c + cv[t][2]) > thr) ci |= 1 << t;
                cube.push_back((uint8_t)ci);
            }
        }
    }

    // ---- a19 mesh: FluidMesher::generateMesh / _generateSurfaceCellMesh /
    // vertexInterp (fluidmesher.cpp:364-545, 4-38), smoothing and normals
    // (fluidmesher.cpp:547-556; trianglemesh.cpp:332-360, 609-630, 736-810) ----
    std::vector<double> mverts;    // xyz per vertex
    std::vector<int32_t> mtris;    // 3 vertex indices per triangle
    std::vector<double> mnormals;  // xyz per vertex
    void mesh(bool smooth) {
        static constexpr McTables T = mc_make_tables();
        const int dd = d();
        const long VI = dd * ni + 1, VJ = dd * nj + 1, VK = dd * nk + 1, NV = VI * VJ * VK;
        const long sni = (long)dd * ni, snj = (long)dd * nj;
        const double sdx = dx / (double)dd;
        const double thr = prm.iso;
        std::vector<int32_t> edge(3 * NV, -1);  // edgeU/V/W keyed by the lower endpoint
        mverts.clear(); mtris.clear();
        for (size_t q = 0; q < surf.size(); ++q) {
            const long a = surf[q] % sni, b = (surf[q] / sni) % snj, c = surf[q] / (sni * snj);
            const int ci = cube[q];
            const uint16_t cut = T.cut[ci];
            if (cut == 0) continue;
            int vl[12];
            for (int e = 0; e < 12; ++e) {
                if (!(cut & (1 << e))) continue;
                const int* l = kMcCorner[kMcEdge[e][0]];
                const int* r = kMcCorner[kMcEdge[e][1]];
                const long li = a + l[0], lj = b + l[1], lk = c + l[2];
                const long key = kMcEdgeDir[e] * NV + li + VI * (lj + VJ * lk);
                if (edge[key] == -1) {
                    const long ri = a + r[0], rj = b + r[1], rk = c + r[2];
                    const double v1 = phi[li + VI * (lj + VJ * lk)], v2 = phi[ri + VI * (rj + VJ * rk)];
                    const double p1[3] = {(li + 0.5) * sdx, (lj + 0.5) * sdx, (lk + 0.5) * sdx};
                    const double p2[3] = {(ri + 0.5) * sdx, (rj + 0.5) * sdx, (rk + 0.5) * sdx};
                    double t = v1 - v2;
                    if (-1e-7 < t && 1e-7 > t) t = 0.5;
                    else t = (v1 - thr) / t;
                    for (int x = 0; x < 3; ++x) mverts.push_back(p1[x] + (p2[x] - p1[x]) * t);
                    edge[key] = (int32_t)(mverts.size() / 3 - 1);
                }
                vl[e] = edge[key];
            }
            for (int t = 0; t < 3 * T.ntri[ci]; ++t) mtris.push_back(vl[T.tri[ci][t]]);
        }
        const long nv = (long)mverts.size() / 3, nt = (long)mtris.size() / 3;
        // _updateVertexTriangles: per vertex, its triangles in ascending order
        std::vector<std::vector<int32_t>> vt(nv);
        for (long t = 0; t < nt; ++t)
            for (int x = 0; x < 3; ++x) vt[mtris[3 * t + x]].push_back((int32_t)t);
        normals(vt);  // generateMesh: updateVertexNormals() before smoothing
        if (!smooth) return;
        // smooth(0.8, 4, all vertices): Jacobi umbrella steps, neighbours with multiplicity
        for (int it = 0; it < 4; ++it) {
            std::vector<double> nvx(3 * nv);
            for (long i = 0; i < nv; ++i) {
                double avg[3] = {0.0, 0.0, 0.0};
                int count = 0;
                for (int32_t t : vt[i])
                    for (int x = 0; x < 3; ++x) {
                        const int32_t o = mtris[3 * t + x];
                        if (o == i) continue;
                        for (int y = 0; y < 3; ++y) avg[y] = avg[y] + mverts[3 * o + y];
                        ++count;
                    }
                for (int y = 0; y < 3; ++y) {
                    const double m = avg[y] / (double)count, v0 = mverts[3 * i + y];
                    nvx[3 * i + y] = v0 + (m - v0) * 0.8;
                }
            }
            mverts.swap(nvx);
        }
        normals(vt);
    }
    // trianglemesh.cpp:332-360 (the sum starts from zero: the reference starts it
    // from a default-constructed, uninitialised Vec3d at :351)
    void normals(const std::vector<std::vector<int32_t>>& vt) {
        const long nv = (long)mverts.size() / 3, nt = (long)mtris.size() / 3;
        std::vector<double> fn(3 * nt);
        auto nrm = [](double* x) {
            const double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            x[0] = x[0] / l; x[1] = x[1] / l; x[2] = x[2] / l;
        };
        for (long t = 0; t < nt; ++t) {
            const double* p0 = &mverts[3 * mtris[3 * t]];
            const double* p1 = &mverts[3 * mtris[3 * t + 1]];
            const double* p2 = &mverts[3 * mtris[3 * t + 2]];
            const double a[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            const double b[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
            nrm(c);
            for (int x = 0; x < 3; ++x) fn[3 * t + x] = c[x];
        }
        mnormals.assign(3 * nv, 0.0);
        for (long i = 0; i < nv; ++i) {
            double n[3] = {0.0, 0.0, 0.0};
            for (int32_t t : vt[i])
                for (int x = 0; x < 3; ++x) n[x] = n[x] + fn[3 * t + x];
            const double cnt = (double)(float)vt[i].size();
            for (int x = 0; x < 3; ++x) n[x] = n[x] / cnt;
            nrm(n);
            for (int x = 0; x < 3; ++x) mnormals[3 * i + x] = n[x];
        }
    }

    // ---- a5: FluidSimulator.cpp:402-427 ----
    void update_material() {
        for (long c = 0; c < N(); ++c) if (mat[c] != SOLID) mat[c] = AIR;
        for (size_t p = 0; p < px.size(); ++p) {
            int i = pos2idx(px[p], dx), j = pos2idx(py[p], dx), k = pos2idx(pz[p], dx);
            if (mat_in(i, j, k) && M(i, j, k) != SOLID) M(i, j, k) = FLUID;
        }
    }

    // ---- a6: FluidSimulator.cpp:1095-1100 ----
    void save_velocity() { us = u; vs = v; ws = w; }

    // ---- a7: FluidSimulator.cpp:429-499 ----
    void apply_forces(double dt) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (size_t f = 0; f < forces.size() / 3; ++f) { ax += forces[3 * f]; ay += forces[3 * f + 1]; az += forces[3 * f + 2]; }
        double dv[3] = {ax * dt, ay * dt, az * dt};
        double err = 1e-10;
        for (int axis = 0; axis < 3; ++axis) {
            if (!(dv[axis] > err || dv[axis] < -err)) continue;
            int si = ni + (axis == 0), sj = nj + (axis == 1), sk = nk + (axis == 2);
            double* f = axis == 0 ? u.data() : axis == 1 ? v.data() : w.data();
            int oi = axis == 0, oj = axis == 1, ok = axis == 2;
            for (int i = 0; i < si; ++i) for (int j = 0; j < sj; ++j) for (int k = 0; k < sk; ++k) {
                bool a = mat_in(i, j, k) && M(i, j, k) == FLUID;
                bool b = mat_in(i - oi, j - oj, k - ok) && M(i - oi, j - oj, k - ok) == FLUID;
                if (a || b) f[i + (long)si * (j + (long)sj * k)] += dv[axis];
            }
        }
    }

    // ---- a8-a14 ----
    int solve() {
        const long n = N();
        std::vector<double> A(7 * n);
        if (prm.pressure_solver == XF_PRESSURE_NORMAL)
            assemble_normal(ni, nj, nk, mat.data(), u.data(), v.data(), w.data(), A.data(), rhs.data());
        else
            assemble(ni, nj, nk, mat.data(), u.data(), v.data(), w.data(), A.data(), rhs.data());
        int it = 0;
        double res = 0;
        int st = pcg(ni, nj, nk, A.data(), rhs.data(), pressure.data(), prm.pcg_tol, prm.pcg_max_iter, prm.mic_tau,
                     prm.mic_sigma, &it, &res, micinv.data());
        last.status = st; last.iterations = it; last.outer = 1; last.residual = res;
        last.rhs_norm = absmax(rhs.data(), n);
        return st;
    }
    void apply() {
        if (prm.pressure_solver == XF_PRESSURE_NORMAL)
            apply_pressure_normal(ni, nj, nk, mat.data(), pressure.data(), u.data(), v.data(), w.data());
        else
            apply_pressure(ni, nj, nk, mat.data(), pressure.data(), u.data(), v.data(), w.data());
    }

    // ---- a15: FluidSimulator.cpp:511-650 ----
    void extrapolate_axis(int layers, int axis) {
        int di = axis == 0, dj = axis == 1, dk = axis == 2;
        int si = ni + di, sj = nj + dj, sk = nk + dk;
        double* f = axis == 0 ? u.data() : axis == 1 ? v.data() : w.data();
        std::vector<int> mark((long)si * sj * sk, -1);
        auto F = [&](int i, int j, int k) { return i + (long)si * (j + (long)sj * k); };
        auto in = [&](int i, int j, int k) { return i >= 0 && j >= 0 && k >= 0 && i < si && j < sj && k < sk; };
        for (int i = 0; i < ni; ++i) for (int j = 0; j < nj; ++j) for (int k = 0; k < nk; ++k)
            if (M(i, j, k) == FLUID) { mark[F(i, j, k)] = 0; mark[F(i + di, j + dj, k + dk)] = 0; }
        const int ndi[6] = {1, -1, 0, 0, 0, 0}, ndj[6] = {0, 0, 1, -1, 0, 0}, ndk[6] = {0, 0, 0, 0, 1, -1};
        std::vector<int> cur, nxt;
        for (int i = 0; i < si; ++i) for (int j = 0; j < sj; ++j) for (int k = 0; k < sk; ++k) {
            if (mark[F(i, j, k)] != -1) continue;
            for (int e = 0; e < 6; ++e) {
                int a = i + ndi[e], b = j + ndj[e], c = k + ndk[e];
                if (in(a, b, c) && mark[F(a, b, c)] == 0) { cur.push_back(i); cur.push_back(j); cur.push_back(k); mark[F(i, j, k)] = 1; break; }
            }
        }
        for (int L = 1; L <= layers; ++L) {
            nxt.clear();
            for (size_t t = 0; t < cur.size(); t += 3) {
                int i = cur[t], j = cur[t + 1], k = cur[t + 2];
                double sum = 0.0;
                int count = 0;
                for (int e = 0; e < 6; ++e) {
                    int a = i + ndi[e], b = j + ndj[e], c = k + ndk[e];
                    if (in(a, b, c) && mark[F(a, b, c)] == L - 1) { sum += f[F(a, b, c)]; count += 1; }
                }
                if (count > 0) {
                    f[F(i, j, k)] = sum / (double)count;
                    for (int e = 0; e < 6; ++e) {
                        int a = i + ndi[e], b = j + ndj[e], c = k + ndk[e];
                        if (in(a, b, c) && mark[F(a, b, c)] == -1) {
                            nxt.push_back(a); nxt.push_back(b); nxt.push_back(c);
                            mark[F(a, b, c)] = L + 1;
                        }
                    }
                }
            }
            cur.swap(nxt);
        }
    }
    // VJP of extrapolate_axis for gradient gf (in place), fixed material.  Layer
    // L = the faces first reached at step L (each the mean of its layer L-1
    // neighbours).  Last layer first, a face of layer L-1 gathers g(t)/count(t)
    // from its layer-L neighbours t (+i,-i,+j,-j,+k,-k); then the faces of
    // layers >= 1 drop their gradient (their values were overwritten).
    void extrapolate_adjoint_axis(int layers, int axis, double* gf) const {
        const int di = axis == 0, dj = axis == 1, dk = axis == 2;
        const int si = ni + di, sj = nj + dj, sk = nk + dk;
        const long nf = (long)si * sj * sk;
        auto F = [&](int i, int j, int k) { return i + (long)si * (j + (long)sj * k); };
        auto in = [&](int i, int j, int k) { return i >= 0 && j >= 0 && k >= 0 && i < si && j < sj && k < sk; };
        std::vector<int> mark(nf, -1);
        for (int i = 0; i < ni; ++i) for (int j = 0; j < nj; ++j) for (int k = 0; k < nk; ++k)
            if (mat[i + (long)ni * (j + (long)nj * k)] == FLUID) { mark[F(i, j, k)] = 0; mark[F(i + di, j + dj, k + dk)] = 0; }
        const int ndi[6] = {1, -1, 0, 0, 0, 0}, ndj[6] = {0, 0, 1, -1, 0, 0}, ndk[6] = {0, 0, 0, 0, 1, -1};
        auto count = [&](int i, int j, int k, int want) {
            int c = 0;
            for (int e = 0; e < 6; ++e)
                if (in(i + ndi[e], j + ndj[e], k + ndk[e]) && mark[F(i + ndi[e], j + ndj[e], k + ndk[e])] == want) ++c;
            return c;
        };
        for (int L = 1; L <= layers; ++L) {
            std::vector<long> reached;
            for (int k = 0; k < sk; ++k) for (int j = 0; j < sj; ++j) for (int i = 0; i < si; ++i)
                if (mark[F(i, j, k)] == -1 && count(i, j, k, L - 1) > 0) reached.push_back(F(i, j, k));
            for (long t : reached) mark[t] = L;
        }
        for (int L = layers; L >= 1; --L)
            for (int k = 0; k < sk; ++k) for (int j = 0; j < sj; ++j) for (int i = 0; i < si; ++i) {
                if (mark[F(i, j, k)] != L - 1) continue;
                double acc = 0.0;
                for (int e = 0; e < 6; ++e) {
                    const int a = i + ndi[e], b = j + ndj[e], c = k + ndk[e];
                    if (!in(a, b, c) || mark[F(a, b, c)] != L) continue;
                    acc += gf[F(a, b, c)] / (double)count(a, b, c, L - 1);
                }
                gf[F(i, j, k)] += acc;
            }
        for (long t = 0; t < nf; ++t)
            if (mark[t] >= 1) gf[t] = 0.0;
    }
    void extrapolate() {
        int layers = (int)(prm.cfl + 5);
        extrapolate_axis(layers, 0); extrapolate_axis(layers, 1); extrapolate_axis(layers, 2);
    }

    // ---- a16: FluidSimulator.cpp:733-894 ----
    void get_velocity(const double* p, double* out, const double* U, const double* Vv, const double* Wv) const {
        if (!((p[0] >= 0.0) && (p[1] >= 0.0) && (p[2] >= 0.0) && (p[0] <= dx * ni) && (p[1] <= dx * nj) &&
              (p[2] <= dx * nk))) {
            out[0] = out[1] = out[2] = 0.0;
            return;
        }
        double invdx = 1.0 / dx;
        double t[8];
        double x, y, z;
        int i, j, k;
        // U: dims (ni+1, nj, nk)
        x = p[0] * invdx; y = (p[1] - 0.5 * dx) * invdx; z = (p[2] - 0.5 * dx) * invdx;
        i = (int)x;
        j = y < 0.0 ? 0 : (int)y; j = j > (nj - 2) ? (nj - 2) : j;
        k = z < 0.0 ? 0 : (int)z; k = k > (nk - 2) ? (nk - 2) : k;
        x -= i; y -= j; z -= k;
        {
            const long si = ni + 1, sj = nj;
            auto G = [&](int a, int b, int c) { return U[a + si * (b + sj * c)]; };
            t[0] = G(i, j, k); t[1] = G(i + 1, j, k); t[2] = G(i, j + 1, k); t[3] = G(i + 1, j + 1, k);
            t[4] = G(i, j, k + 1); t[5] = G(i + 1, j, k + 1); t[6] = G(i, j + 1, k + 1); t[7] = G(i + 1, j + 1, k + 1);
        }
        out[0] = trilerp(t, x, y, z);
        // V: dims (ni, nj+1, nk)
        x = (p[0] - 0.5 * dx) * invdx; y = p[1] * invdx; z = (p[2] - 0.5 * dx) * invdx;
        j = (int)y;
        i = x < 0.0 ? 0 : (int)x; i = i > (ni - 2) ? (ni - 2) : i;
        k = z < 0.0 ? 0 : (int)z; k = k > (nk - 2) ? (nk - 2) : k;
        x -= i; y -= j; z -= k;
        {
            const long si = ni, sj = nj + 1;
            auto G = [&](int a, int b, int c) { return Vv[a + si * (b + sj * c)]; };
            t[0] = G(i, j, k); t[1] = G(i + 1, j, k); t[2] = G(i, j + 1, k); t[3] = G(i + 1, j + 1, k);
            t[4] = G(i, j, k + 1); t[5] = G(i + 1, j, k + 1); t[6] = G(i, j + 1, k + 1); t[7] = G(i + 1, j + 1, k + 1);
        }
        out[1] = trilerp(t, x, y, z);
        // W: dims (ni, nj, nk+1)
        z = p[2] * invdx; y = (p[1] - 0.5 * dx) * invdx; x = (p[0] - 0.5 * dx) * invdx;
        k = (int)z;
        j = y < 0.0 ? 0 : (int)y; j = j > (nj - 2) ? (nj - 2) : j;
        i = x < 0.0 ? 0 : (int)x; i = i > (ni - 2) ? (ni - 2) : i;
        x -= i; y -= j; z -= k;
        {
            const long si = ni, sj = nj;
            auto G = [&](int a, int b, int c) { return Wv[a + si * (b + sj * c)]; };
            t[0] = G(i, j, k); t[1] = G(i + 1, j, k); t[2] = G(i, j + 1, k); t[3] = G(i + 1, j + 1, k);
            t[4] = G(i, j, k + 1); t[5] = G(i + 1, j, k + 1); t[6] = G(i, j + 1, k + 1); t[7] = G(i + 1, j + 1, k + 1);
        }
        out[2] = trilerp(t, x, y, z);
    }

    // FluidSimulator.cpp:896-907 (RK4); RK2 midpoint variant
    void advect_point(const double* p, double dt, double* out) const {
        double k1[3], k2[3], k3[3], k4[3], q[3];
        const double* U = u.data(); const double* V = v.data(); const double* W = w.data();
        get_velocity(p, k1, U, V, W);
        if (prm.integrator == XF_INTEGRATOR_RK2) {
            for (int a = 0; a < 3; ++a) q[a] = p[a] + k1[a] * (0.5 * dt);
            get_velocity(q, k2, U, V, W);
            for (int a = 0; a < 3; ++a) out[a] = p[a] + k2[a] * dt;
            return;
        }
        for (int a = 0; a < 3; ++a) q[a] = p[a] + k1[a] * (0.5 * dt);
        get_velocity(q, k2, U, V, W);
        for (int a = 0; a < 3; ++a) q[a] = p[a] + k2[a] * (0.5 * dt);
        get_velocity(q, k3, U, V, W);
        for (int a = 0; a < 3; ++a) q[a] = p[a] + k3[a] * dt;
        get_velocity(q, k4, U, V, W);
        for (int a = 0; a < 3; ++a) out[a] = p[a] + (k1[a] + k2[a] * 2.0 + k3[a] * 2.0 + k4[a]) * (dt / 6.0);
    }

    // collision.cpp:293-385
    bool voxel_hit(const double* a, const double* b, int* vox) const {
        double invdx = 1.0 / dx;
        double p0[3] = {a[0] * invdx, a[1] * invdx, a[2] * invdx};
        double p1[3] = {b[0] * invdx, b[1] * invdx, b[2] * invdx};
        int g0[3], g1[3], sd[3], g[3];
        double gp[3], vv[3];
        for (int t = 0; t < 3; ++t) {
            g0[t] = (int)std::floor(p0[t]); g1[t] = (int)std::floor(p1[t]);
            sd[t] = g1[t] > g0[t] ? 1 : g1[t] < g0[t] ? -1 : 0;
            g[t] = g0[t];
            gp[t] = g0[t] + (g1[t] > g0[t] ? 1 : 0);
            vv[t] = p1[t] == p0[t] ? 1 : p1[t] - p0[t];
        }
        double vxvy = vv[0] * vv[1], vxvz = vv[0] * vv[2], vyvz = vv[1] * vv[2];
        double errx = (gp[0] - p0[0]) * vyvz, erry = (gp[1] - p0[1]) * vxvz, errz = (gp[2] - p0[2]) * vxvy;
        double derrx = sd[0] * vyvz, derry = sd[1] * vxvz, derrz = sd[2] * vxvy;
        for (;;) {
            if (g[0] >= 0 && g[1] >= 0 && g[2] >= 0 && g[0] < ni && g[1] < nj && g[2] < nk &&
                mat[g[0] + (long)ni * (g[1] + (long)nj * g[2])] == SOLID) {
                vox[0] = g[0]; vox[1] = g[1]; vox[2] = g[2];
                return true;
            }
            if (g[0] == g1[0] && g[1] == g1[1] && g[2] == g1[2]) break;
            double xr = std::fabs(errx), yr = std::fabs(erry), zr = std::fabs(errz);
            if (sd[0] != 0 && (sd[1] == 0 || xr < yr) && (sd[2] == 0 || xr < zr)) { g[0] += sd[0]; errx += derrx; }
            else if (sd[1] != 0 && (sd[2] == 0 || yr < zr)) { g[1] += sd[1]; erry += derry; }
            else if (sd[2] != 0) { g[2] += sd[2]; errz += derrz; }
        }
        return false;
    }

    // collision.cpp:396-439, with the reference's repeated dir.x guard; also
    // the axis of the face the ray enters through and whether it is the upper one
    static bool ray_box(const double* p0, const double* dir_in, const double* bmin, const double* bmax, double* hit,
                        int* axis = nullptr, bool* upper = nullptr) {
        double dir[3] = {dir_in[0], dir_in[1], dir_in[2]};
        double eps = 1e-10;
        if (std::fabs(dir[0]) < eps) dir[0] = dir[0] < 0 ? -eps : eps;
        if (std::fabs(dir[1]) < eps) dir[1] = dir[1] < 0 ? -eps : eps;
        if (std::fabs(dir[0]) < eps) dir[2] = dir[2] < 0 ? -eps : eps;
        double inv[3] = {1.0 / dir[0], 1.0 / dir[1], 1.0 / dir[2]};
        double t1 = (bmin[0] - p0[0]) * inv[0], t2 = (bmax[0] - p0[0]) * inv[0];
        double tmin = std::fmin(t1, t2), tmax = std::fmax(t1, t2);
        t1 = (bmin[1] - p0[1]) * inv[1]; t2 = (bmax[1] - p0[1]) * inv[1];
        const double m1 = std::fmin(t1, t2);
        tmin = std::fmax(tmin, m1); tmax = std::fmin(tmax, std::fmax(t1, t2));
        t1 = (bmin[2] - p0[2]) * inv[2]; t2 = (bmax[2] - p0[2]) * inv[2];
        const double m2 = std::fmin(t1, t2);
        tmin = std::fmax(tmin, m2); tmax = std::fmin(tmax, std::fmax(t1, t2));
        if (tmax > std::fmax(tmin, 0.0)) {
            for (int a = 0; a < 3; ++a) hit[a] = p0[a] + dir[a] * tmin;
            if (axis) {
                *axis = tmin == m2 ? 2 : tmin == m1 ? 1 : 0;
                *upper = inv[*axis] < 0.0;
            }
            return true;
        }
        return false;
    }

    // the end of a particle's move (FluidSimulator.cpp:686-722): 0 the
    // integrated target, 1 the old position kept, 2 pulled back from the first
    // solid voxel on the way; for 2, `c` holds the ray, the entry face's axis
    // and its distance (what the adjoint differentiates)
    struct Coll {
        double rn[3], l, tmin, back;
        int axis;
    };
    int resolve_move(const double* pos, const double* np_, double* out, Coll* c) const {
        int gi = pos2idx(np_[0], dx), gj = pos2idx(np_[1], dx), gk = pos2idx(np_[2], dx);
        if (mat_in(gi, gj, gk) && mat[gi + (long)ni * (gj + (long)nj * gk)] != SOLID) {
            out[0] = np_[0]; out[1] = np_[1]; out[2] = np_[2];
            return 0;
        }
        int vox[3];
        if (!voxel_hit(pos, np_, vox)) return 1;
        double dv[3] = {np_[0] - pos[0], np_[1] - pos[1], np_[2] - pos[2]};
        double l = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        double rn[3] = {dv[0] / l, dv[1] / l, dv[2] / l};
        double bmin[3] = {idx2pos(vox[0], dx), idx2pos(vox[1], dx), idx2pos(vox[2], dx)};
        double bmax[3] = {bmin[0] + dx, bmin[1] + dx, bmin[2] + dx};
        double cp[3];
        int axis = 0;
        bool upper = false;
        if (!ray_box(pos, rn, bmin, bmax, cp, &axis, &upper)) return 1;
        double back = (double)(float)(0.05 * dx);
        double res[3] = {cp[0] - rn[0] * back, cp[1] - rn[1] * back, cp[2] - rn[2] * back};
        int ri = pos2idx(res[0], dx), rj = pos2idx(res[1], dx), rk = pos2idx(res[2], dx);
        // the reference indexes material without a range check; out of range is rejected here
        if (!(mat_in(ri, rj, rk) && mat[ri + (long)ni * (rj + (long)nj * rk)] != SOLID)) return 1;
        out[0] = res[0]; out[1] = res[1]; out[2] = res[2];
        for (int a = 0; a < 3; ++a) c->rn[a] = rn[a];
        c->l = l;
        c->axis = axis;
        c->tmin = ((upper ? bmax[axis] : bmin[axis]) - pos[axis]) / rn[axis];
        c->back = back;
        return 2;
    }

    // FluidSimulator.cpp:652-723 (G2P blend, RK4 advection, collision)
    // FluidSimulator::addFluidSource (FluidSimulator.cpp:152-166) +
    // CubicFluidSource::setCubic (fluidsource.cpp:8-33) and setVelocity (:35-39):
    // the cells between pos2Index(c - w) and pos2Index(c + w), i slowest, in range
    // and not solid now
    void add_fluid_source(const double* c, double wx, double wy, double wz, const double* vel) {
        Source src;
        const double lo[3] = {c[0] - wx, c[1] - wy, c[2] - wz}, hi[3] = {c[0] + wx, c[1] + wy, c[2] + wz};
        const int gi0 = (int)(lo[0] / dx), gj0 = (int)(lo[1] / dx), gk0 = (int)(lo[2] / dx);
        const int gi1 = (int)(hi[0] / dx), gj1 = (int)(hi[1] / dx), gk1 = (int)(hi[2] / dx);
        for (int i = gi0; i <= gi1; ++i)
            for (int j = gj0; j <= gj1; ++j)
                for (int k = gk0; k <= gk1; ++k)
                    if (mat_in(i, j, k) && M(i, j, k) != SOLID) src.cells.push_back(i + (long)ni * j + (long)ni * nj * k);
        for (int a = 0; a < 3; ++a) src.v[a] = vel[a];
        src.normv = std::sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
        sources.push_back(src);
    }
    // randomDouble (xfluidutil.cpp:67-70) over a 31-bit draw, RAND_MAX = 2^31 - 1
    static double random_double(double lo, double hi, uint32_t r) {
        return lo + (double)r / ((double)2147483647 / (hi - lo));
    }
    // _applyFluidSource (FluidSimulator.cpp:1087-1093) -> CubicFluidSource::update
    // (fluidsource.cpp:75-92).  Draws x, y, z per particle from the stream, or from
    // `draws` (three per particle, in that order) when given.
    int apply_sources(double dt, const uint32_t* draws = nullptr) {
        long d = 0;
        for (const Source& src : sources) {
            const double x = 8.0 * src.normv * dt / dx;  // _particlePerCell * _normv * dt / _dx
            if (!(x < 2147483647.0)) return -1;
            const int nnew = x > 0 ? (int)x : 0;
            for (long c : src.cells) {
                mat[c] = FLUID;
                const int i = (int)(c % ni), j = (int)((c / ni) % nj), k = (int)(c / ((long)ni * nj));
                for (int q = 0; q < nnew; ++q) {
                    double p[3] = {(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx};  // idx2Pos
                    for (int a = 0; a < 3; ++a) {
                        const uint32_t r = draws ? draws[d++] : (uint32_t)(splitmix64(src_state) >> 33);
                        p[a] = p[a] + random_double(-0.4999 * dx, 0.4999 * dx, r);
                    }
                    px.push_back(p[0]); py.push_back(p[1]); pz.push_back(p[2]);
                    vx.push_back(src.v[0]); vy.push_back(src.v[1]); vz.push_back(src.v[2]);
                }
            }
        }
        return 0;
    }

    void advect(double dt) {
        const long n = (long)px.size();
        const long lim = prm.tail_skip ? (n / 6) * 6 : n;
        const double r = prm.flip_ratio;
#pragma omp parallel for schedule(static)
        for (long p = 0; p < lim; ++p) {
            double pos[3] = {px[p], py[p], pz[p]};
            double vnew[3], vold[3];
            get_velocity(pos, vnew, u.data(), v.data(), w.data());
            get_velocity(pos, vold, us.data(), vs.data(), ws.data());
            double vp[3] = {vx[p], vy[p], vz[p]};
            for (int a = 0; a < 3; ++a) {
                double vpic = vnew[a];
                double vflip = vp[a] + vnew[a] - vold[a];
                vp[a] = vpic * (1.0 - r) + vflip * r;
            }
            vx[p] = vp[0]; vy[p] = vp[1]; vz[p] = vp[2];

            double np_[3], out[3];
            advect_point(pos, dt, np_);
            Coll c;
            if (resolve_move(pos, np_, out, &c) != 1) {
                px[p] = out[0]; py[p] = out[1]; pz[p] = out[2];
            }
        }
    }

    // ---- a18: FluidSimulator.cpp:909-1085 (Wyvill r = 1.5dx, normalised) ----
    // Sums in particle order over each particle's face box (the reference's
    // loops), the weight from p2g_visit.
    void p2g_axis(int axis) {
        const int si = ni + (axis == 0), sj = nj + (axis == 1), sk = nk + (axis == 2);
        const long nf = (long)si * sj * sk;
        std::vector<double> wt(nf, 0.0), val(nf, 0.0);
        const double err = 0.0000001;
        const double* vc = axis == 0 ? vx.data() : axis == 1 ? vy.data() : vz.data();
        for (size_t p = 0; p < px.size(); ++p) {
            const double pp[3] = {px[p], py[p], pz[p]};
            const double vv = vc[p];
            p2g_visit<false>(axis, pp, [&](long f, double cw, const double*) {
                wt[f] += cw;
                val[f] += vv * cw;
            });
        }
        double* out = axis == 0 ? u.data() : axis == 1 ? v.data() : w.data();
        for (long f = 0; f < nf; ++f) out[f] = wt[f] > err ? val[f] / wt[f] : 0.0;
    }
    void p2g() { p2g_axis(0); p2g_axis(1); p2g_axis(2); }

    // ---- adjoints of the particle transfers (no reference counterpart; the
    // reference has no gradients).  Particles are passed as n x 3 arrays. ----

    // P2G of one axis for given particles: weight sums wt and values out (as p2g_axis)
    void p2g_points(int axis, long n, const double* pos, const double* vel, double* out, double* wt) const {
        const int si = ni + (axis == 0), sj = nj + (axis == 1), sk = nk + (axis == 2);
        const long nf = (long)si * sj * sk;
        std::vector<double> val(nf, 0.0);
        std::fill(wt, wt + nf, 0.0);
        for (long p = 0; p < n; ++p)
            p2g_visit<false>(axis, pos + 3 * p, [&](long f, double cw, const double*) {
                wt[f] += cw;
                val[f] += vel[3 * p + axis] * cw;
            });
        for (long f = 0; f < nf; ++f) out[f] = wt[f] > 0.0000001 ? val[f] / wt[f] : 0.0;
    }
    // visits the faces of one axis in a particle's support with the weight
    // (GRAD: also the weight's gradient with respect to the particle position)
    template <bool GRAD, class F>
    void p2g_visit(int axis, const double* pp, F&& fn) const {
        const int si = ni + (axis == 0), sj = nj + (axis == 1), sk = nk + (axis == 2);
        const bool tri = prm.p2g_kernel == XF_P2G_TRILINEAR;
        const double r = tri ? dx : 1.5 * dx;
        const double r2 = r * r;
        const double c0 = (4.0 / 9.0) * (1.0 / (r2 * r2 * r2));
        const double c1 = (17.0 / 9.0) * (1.0 / (r2 * r2));
        const double c2 = (22.0 / 9.0) * (1.0 / (r2));
        double q[3] = {pp[0] + (axis == 0 ? 0.5 * dx : 0.0), pp[1] + (axis == 1 ? 0.5 * dx : 0.0),
                       pp[2] + (axis == 2 ? 0.5 * dx : 0.0)};
        int g0[3], g1[3];
        for (int a = 0; a < 3; ++a) { g0[a] = pos2idx(-r + q[a], dx); g1[a] = pos2idx(q[a] + r, dx); }
        for (int i = g0[0]; i <= g1[0]; ++i)
            for (int j = g0[1]; j <= g1[1]; ++j)
                for (int k = g0[2]; k <= g1[2]; ++k) {
                    if (i < 0 || j < 0 || k < 0 || i >= si || j >= sj || k >= sk) continue;
                    double e[3] = {q[0] - idx2pos(i, dx), q[1] - idx2pos(j, dx), q[2] - idx2pos(k, dx)};
                    double cw = 0.0, gw[3] = {0.0, 0.0, 0.0};
                    if (tri) {
                        double a[3] = {std::fabs(e[0]) / dx, std::fabs(e[1]) / dx, std::fabs(e[2]) / dx};
                        if (a[0] < 1.0 && a[1] < 1.0 && a[2] < 1.0) {
                            cw = (1.0 - a[0]) * (1.0 - a[1]) * (1.0 - a[2]);
                            if (GRAD)
                            for (int t = 0; t < 3; ++t) {
                                double sg = e[t] > 0 ? 1.0 : e[t] < 0 ? -1.0 : 0.0;
                                gw[t] = -sg / dx * (1.0 - a[(t + 1) % 3]) * (1.0 - a[(t + 2) % 3]);
                            }
                        }
                    } else {
                        double rs = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                        if (rs < r2) {
                            cw = 1.0 - c0 * rs * rs * rs + c1 * rs * rs - c2 * rs;
                            if (GRAD) {
                                const double dw = -3.0 * c0 * rs * rs + 2.0 * c1 * rs - c2;
                                for (int t = 0; t < 3; ++t) gw[t] = dw * 2.0 * e[t];
                            }
                        }
                    }
                    fn(i + (long)si * (j + (long)sj * k), cw, gw);
                }
    }
    // VJP of P2G: g on the face arrays -> gvel (n x 3, overwritten), gpos (n x 3, accumulated)
    // VJP of density() for the particles pos (n x 3): gpos += sum_v gphi(v) W'(s)
    // (-2)(x_v - p) over each particle's vertex box (directeddistancefield.cpp:
    // 300-328), W'(s) = -3 c0 s^2 + 2 c1 s - c2 the derivative of the Wyvill
    // polynomial (directeddistancefield.cpp:246-256); test infrastructure only.
    void density_adjoint(long n, const double* pos, const double* gphi, double* gpos) {
        const int dd = d();
        const int si = dd * ni + 1, sj = dd * nj + 1, sk = dd * nk + 1;
        const double R = blob_radius(), R2 = R * R;
        const double c0 = (4.0 / 9.0) * (1.0 / (R2 * R2 * R2));
        const double c1 = (17.0 / 9.0) * (1.0 / (R2 * R2));
        const double c2 = (22.0 / 9.0) * (1.0 / (R2));
        const double subdx = dx / (double)dd;
        for (long p = 0; p < n; ++p) {
            const double x = pos[3 * p], y = pos[3 * p + 1], z = pos[3 * p + 2];
            int g0i = pos2idx(-R + x, dx) * dd, g0j = pos2idx(-R + y, dx) * dd, g0k = pos2idx(-R + z, dx) * dd;
            int g1i = (pos2idx(x + R, dx) + 1) * dd, g1j = (pos2idx(y + R, dx) + 1) * dd,
                g1k = (pos2idx(z + R, dx) + 1) * dd;
            double ax = 0.0, ay = 0.0, az = 0.0;
            for (int i = g0i; i <= g1i; ++i)
                for (int j = g0j; j <= g1j; ++j)
                    for (int k = g0k; k <= g1k; ++k) {
                        if (i < 0 || j < 0 || k < 0 || i >= si || j >= sj || k >= sk) continue;
                        double ex = idx2pos(i, subdx) - x, ey = idx2pos(j, subdx) - y, ez = idx2pos(k, subdx) - z;
                        double rs = ex * ex + ey * ey + ez * ez;
                        if (!(rs < R2)) continue;
                        double dw = -3.0 * c0 * rs * rs + 2.0 * c1 * rs - c2;
                        double f = gphi[i + (long)si * (j + (long)sj * k)] * dw * -2.0;
                        ax += f * ex;
                        ay += f * ey;
                        az += f * ez;
                    }
            gpos[3 * p] += ax;
            gpos[3 * p + 1] += ay;
            gpos[3 * p + 2] += az;
        }
    }
    void p2g_adjoint(long n, const double* pos, const double* vel, const double* const g[3], double* gvel,
                     double* gpos) const {
        std::fill(gvel, gvel + 3 * n, 0.0);
        for (int axis = 0; axis < 3; ++axis) {
            const long nf = (long)(ni + (axis == 0)) * (nj + (axis == 1)) * (nk + (axis == 2));
            std::vector<double> uf(nf), wt(nf), a(nf);
            p2g_points(axis, n, pos, vel, uf.data(), wt.data());
            for (long f = 0; f < nf; ++f) a[f] = wt[f] > 0.0000001 ? g[axis][f] / wt[f] : 0.0;
            for (long p = 0; p < n; ++p) {
                const double vp = vel[3 * p + axis];
                p2g_visit<true>(axis, pos + 3 * p, [&](long f, double cw, const double* gw) {
                    if (a[f] == 0.0) return;
                    gvel[3 * p + axis] += a[f] * cw;
                    for (int t = 0; t < 3; ++t) gpos[3 * p + t] += a[f] * (vp - uf[f]) * gw[t];
                });
            }
        }
    }

    // VJP of trilerp (xfluidutil.cpp:53-65): corner weights and d/d(x,y,z)
    static void trilerp_adj(const double* t, double x, double y, double z, double go, double* gt, double* gx) {
        const double wx[2] = {1.0 - x, x}, wy[2] = {1.0 - y, y}, wz[2] = {1.0 - z, z};
        gx[0] = gx[1] = gx[2] = 0.0;
        for (int c = 0; c < 2; ++c)
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    const int q = a + 2 * b + 4 * c;
                    gt[q] = go * wx[a] * wy[b] * wz[c];
                    gx[0] += t[q] * (a ? 1.0 : -1.0) * wy[b] * wz[c];
                    gx[1] += t[q] * wx[a] * (b ? 1.0 : -1.0) * wz[c];
                    gx[2] += t[q] * wx[a] * wy[b] * (c ? 1.0 : -1.0);
                }
        gx[0] *= go; gx[1] *= go; gx[2] *= go;
    }
    // VJP of get_velocity at p: gU/gV/gW and gp accumulated
    void sample_adjoint(const double* p, const double* go, const double* const F[3], double* const gF[3],
                        double* gp) const {
        if (!((p[0] >= 0.0) && (p[1] >= 0.0) && (p[2] >= 0.0) && (p[0] <= dx * ni) && (p[1] <= dx * nj) &&
              (p[2] <= dx * nk)))
            return;
        const double invdx = 1.0 / dx;
        for (int comp = 0; comp < 3; ++comp) {
            double x, y, z;
            int i, j, k;
            long si, sj;
            if (comp == 0) {
                x = p[0] * invdx; y = (p[1] - 0.5 * dx) * invdx; z = (p[2] - 0.5 * dx) * invdx;
                i = (int)x;
                j = y < 0.0 ? 0 : (int)y; j = j > (nj - 2) ? (nj - 2) : j;
                k = z < 0.0 ? 0 : (int)z; k = k > (nk - 2) ? (nk - 2) : k;
                si = ni + 1; sj = nj;
            } else if (comp == 1) {
                x = (p[0] - 0.5 * dx) * invdx; y = p[1] * invdx; z = (p[2] - 0.5 * dx) * invdx;
                j = (int)y;
                i = x < 0.0 ? 0 : (int)x; i = i > (ni - 2) ? (ni - 2) : i;
                k = z < 0.0 ? 0 : (int)z; k = k > (nk - 2) ? (nk - 2) : k;
                si = ni; sj = nj + 1;
            } else {
                z = p[2] * invdx; y = (p[1] - 0.5 * dx) * invdx; x = (p[0] - 0.5 * dx) * invdx;
                k = (int)z;
                j = y < 0.0 ? 0 : (int)y; j = j > (nj - 2) ? (nj - 2) : j;
                i = x < 0.0 ? 0 : (int)x; i = i > (ni - 2) ? (ni - 2) : i;
                si = ni; sj = nj;
            }
            x -= i; y -= j; z -= k;
            long idx[8];
            double t[8], gt[8], gx[3];
            for (int q = 0; q < 8; ++q) {
                idx[q] = (i + (q & 1)) + si * ((j + ((q >> 1) & 1)) + sj * (k + (q >> 2)));
                t[q] = F[comp][idx[q]];
            }
            trilerp_adj(t, x, y, z, go[comp], gt, gx);
            for (int q = 0; q < 8; ++q) gF[comp][idx[q]] += gt[q];
            for (int a = 0; a < 3; ++a) gp[a] += gx[a] * invdx;
        }
    }
    // VJP of advect_point (RK4 / RK2, collision-free): gnew -> gp (overwritten), gF accumulated
    // VJP of the whole position update p -> resolve_move(p, RK(p)) for gradient
    // gout on the new position: the gradient reaching the integrated target
    // (grk) and the part that goes to p directly (gextra).  Kept: all to p.
    // Pulled back (X' = p + rn (t - back), t = (B - p_a) / rn_a, rn = d / |d|,
    // d = target - p):  g_rn = (t - back) g - e_a (g.rn) t / rn_a,
    // g_d = (I - rn rn^T) g_rn / |d|,  grk = g_d,  gextra = g - e_a (g.rn) / rn_a - g_d.
    void move_adjoint(const double* p, const double* np_, const double* gout, double* grk, double* gextra) const {
        double out[3];
        Coll c;
        const int kind = resolve_move(p, np_, out, &c);
        for (int a = 0; a < 3; ++a) { grk[a] = kind == 0 ? gout[a] : 0.0; gextra[a] = kind == 1 ? gout[a] : 0.0; }
        if (kind != 2) return;
        const int ax = c.axis;
        const double s = gout[0] * c.rn[0] + gout[1] * c.rn[1] + gout[2] * c.rn[2];
        const double cc = c.tmin - c.back;
        double grn[3];
        for (int k = 0; k < 3; ++k) grn[k] = cc * gout[k] - (k == ax ? s * c.tmin / c.rn[ax] : 0.0);
        const double rg = c.rn[0] * grn[0] + c.rn[1] * grn[1] + c.rn[2] * grn[2];
        for (int k = 0; k < 3; ++k) {
            grk[k] = (grn[k] - c.rn[k] * rg) / c.l;
            gextra[k] = gout[k] - (k == ax ? s / c.rn[ax] : 0.0) - grk[k];
        }
    }
    // the integrated target of advect_point followed by resolve_move
    void move_point(const double* p, double dt, double* out) const {
        double np_[3];
        advect_point(p, dt, np_);
        Coll c;
        if (resolve_move(p, np_, out, &c) == 1) { out[0] = p[0]; out[1] = p[1]; out[2] = p[2]; }
    }
    // VJP of move_point: gnew on the new position -> gp (written), gF (added)
    void move_adjoint_full(const double* p, double dt, const double* gnew, double* const gF[3], double* gp) const {
        double np_[3], grk[3], gextra[3];
        advect_point(p, dt, np_);
        move_adjoint(p, np_, gnew, grk, gextra);
        advect_adjoint(p, dt, grk, gF, gp);
        for (int a = 0; a < 3; ++a) gp[a] += gextra[a];
    }
    void advect_adjoint(const double* p, double dt, const double* gnew, double* const gF[3], double* gp) const {
        const double* F[3] = {u.data(), v.data(), w.data()};
        double k1[3], k2[3], k3[3], q2[3], q3[3], q4[3], g[3];
        for (int a = 0; a < 3; ++a) gp[a] = gnew[a];
        get_velocity(p, k1, F[0], F[1], F[2]);
        if (prm.integrator == XF_INTEGRATOR_RK2) {
            for (int a = 0; a < 3; ++a) q2[a] = p[a] + k1[a] * (0.5 * dt);
            double gk2[3] = {gnew[0] * dt, gnew[1] * dt, gnew[2] * dt}, gq2[3] = {0, 0, 0};
            sample_adjoint(q2, gk2, F, gF, gq2);
            double gk1[3];
            for (int a = 0; a < 3; ++a) { gp[a] += gq2[a]; gk1[a] = gq2[a] * (0.5 * dt); }
            sample_adjoint(p, gk1, F, gF, gp);
            return;
        }
        for (int a = 0; a < 3; ++a) q2[a] = p[a] + k1[a] * (0.5 * dt);
        get_velocity(q2, k2, F[0], F[1], F[2]);
        for (int a = 0; a < 3; ++a) q3[a] = p[a] + k2[a] * (0.5 * dt);
        get_velocity(q3, k3, F[0], F[1], F[2]);
        for (int a = 0; a < 3; ++a) q4[a] = p[a] + k3[a] * dt;
        double gk1[3], gk2[3], gk3[3], gk4[3];
        for (int a = 0; a < 3; ++a) {
            g[a] = gnew[a] * (dt / 6.0);
            gk1[a] = g[a]; gk2[a] = g[a] * 2.0; gk3[a] = g[a] * 2.0; gk4[a] = g[a];
        }
        double gq[3] = {0, 0, 0};
        sample_adjoint(q4, gk4, F, gF, gq);
        for (int a = 0; a < 3; ++a) { gp[a] += gq[a]; gk3[a] += gq[a] * dt; gq[a] = 0; }
        sample_adjoint(q3, gk3, F, gF, gq);
        for (int a = 0; a < 3; ++a) { gp[a] += gq[a]; gk2[a] += gq[a] * (0.5 * dt); gq[a] = 0; }
        sample_adjoint(q2, gk2, F, gF, gq);
        for (int a = 0; a < 3; ++a) { gp[a] += gq[a]; gk1[a] += gq[a] * (0.5 * dt); }
        sample_adjoint(p, gk1, F, gF, gp);
    }

    int stage(int s, double dt) {
        switch (s) {
            case XF_STAGE_DENSITY: density(); surface(); return 0;
            case XF_STAGE_MATERIAL: update_material(); return 0;
            case XF_STAGE_SAVE: save_velocity(); return 0;
            case XF_STAGE_FORCES: apply_forces(dt); return 0;
            case XF_STAGE_SOLVE: return solve();
            case XF_STAGE_APPLY: apply(); return 0;
            case XF_STAGE_EXTRAPOLATE: extrapolate(); return 0;
            case XF_STAGE_ADVECT: advect(dt); return apply_sources(dt);
            case XF_STAGE_P2G: p2g(); return 0;
        }
        return -1;
    }

    // FluidSimulator.cpp:265-335
    int time_step(double dt, bool mesh) {
        if (prm.density_every_substep || mesh) { density(); if (mesh) surface(); }
        update_material();
        save_velocity();
        apply_forces(dt);
        int st = solve();
        apply();
        extrapolate();
        advect(dt);
        if (apply_sources(dt) < 0) return -1;
        p2g();
        return st;
    }

    // FluidSimulator.cpp:224-262, 387-400
    int frame_step(double dt) {
        double left = dt;
        int worst = 0;
        while (left > 0.0) {
            double maxv = 0.0;
            for (size_t p = 0; p < vx.size(); ++p) {
                double c = vx[p] * vx[p] + vy[p] * vy[p] + vz[p] * vz[p];
                maxv = c > maxv ? c : maxv;
            }
            maxv = std::sqrt(maxv);
            double cur = maxv == 0.0 ? prm.max_dt : prm.cfl * dx / maxv;
            cur = cur < prm.max_dt ? cur : prm.max_dt;
            if (left <= cur) { cur = left; left = 0.0; } else left -= cur;
            int st = time_step(cur, left <= 0.0);
            if (st > worst) worst = st;
        }
        ++frame;
        return worst;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

void xf_params_default(xf_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->flip_ratio = 0.95; p->cfl = 1.0; p->max_dt = 0.5; p->pcg_tol = 1e-5; p->pcg_max_iter = 100;
    p->sub_division = 2; p->mic_tau = 0.97; p->mic_sigma = 0.25; p->blob_radius = 0.0; p->iso = 0.4;
    p->precision = XF_PREC_FP64; p->integrator = XF_INTEGRATOR_RK4; p->p2g_kernel = XF_P2G_WYVILL;
    p->tail_skip = 1; p->density_every_substep = 1; p->sweep_kernel = 0;
    p->mixed_inner_tol = 1e-3; p->mixed_max_outer = 20;
}

// -- standalone pressure pieces (A as 7 planes, see enum above) --
void xo_assemble(int ni, int nj, int nk, const uint8_t* mat, const double* u, const double* v, const double* w,
                 double* A7, double* b) { assemble(ni, nj, nk, mat, u, v, w, A7, b); }
void xo_spmv(int ni, int nj, int nk, const double* A7, const double* x, double* y) { spmv(ni, nj, nk, A7, x, y); }
void xo_mic0(int ni, int nj, int nk, const double* A7, double tau, double sigma, double* inv) {
    mic0(ni, nj, nk, A7, tau, sigma, inv);
}
void xo_sweep_fwd(int ni, int nj, int nk, const double* A7, const double* inv, const double* r, double* q) {
    sweep_fwd(ni, nj, nk, A7, inv, r, q);
}
void xo_sweep_bwd(int ni, int nj, int nk, const double* A7, const double* inv, const double* q, double* z) {
    long N = (long)ni * nj * nk;
    std::copy(q, q + N, z);
    sweep_bwd(ni, nj, nk, A7, inv, z);
}
int xo_pcg(int ni, int nj, int nk, const double* A7, const double* b, double* x, double tol, int max_iter,
           double tau, double sigma, int* iters, double* res) {
    return pcg(ni, nj, nk, A7, b, x, tol, max_iter, tau, sigma, iters, res, nullptr);
}
void xo_apply_pressure(int ni, int nj, int nk, const uint8_t* mat, const double* p, double* u, double* v, double* w) {
    apply_pressure(ni, nj, nk, mat, p, u, v, w);
}
void xo_assemble_normal(int ni, int nj, int nk, const uint8_t* mat, const double* u, const double* v,
                        const double* w, double* A7, double* b) {
    assemble_normal(ni, nj, nk, mat, u, v, w, A7, b);
}
void xo_apply_pressure_normal(int ni, int nj, int nk, const uint8_t* mat, const double* p, double* u, double* v,
                              double* w) {
    apply_pressure_normal(ni, nj, nk, mat, p, u, v, w);
}

// -- simulator handle --
void* xo_create(int ni, int nj, int nk) {
    if (ni < 3 || nj < 3 || nk < 3) return nullptr;
    return new Sim(ni, nj, nk);
}
void xo_destroy(void* h) { delete (Sim*)h; }
void xo_set_cell_width(void* h, double dx) { ((Sim*)h)->dx = dx; }
void xo_set_params(void* h, const xf_params* p) {
    Sim* s = (Sim*)h;
    int old = s->prm.sub_division;
    s->prm = *p;
    if (s->prm.sub_division != old) s->phi.assign(s->nphi(), 0.0);
}
void xo_set_solid_boundary(void* h) { ((Sim*)h)->set_solid_boundary(); }
void xo_add_sphere_fluid(void* h, double x, double y, double z, double r) { ((Sim*)h)->add_sphere(x, y, z, r); }
void xo_add_cubic_fluid(void* h, const double* lu, const double* rd) { ((Sim*)h)->add_cubic(lu, rd); }
void xo_add_constant_force(void* h, double fx, double fy, double fz) {
    Sim* s = (Sim*)h;
    s->forces.push_back(fx); s->forces.push_back(fy); s->forces.push_back(fz);
}
void xo_seed_particles(void* h, int mode, uint64_t seed) { ((Sim*)h)->seed(mode, seed); }
// reference-mode seeding with injected 31-bit draws (3 per particle, x, y, z)
void xo_seed_particles_draws(void* h, const uint32_t* draws) { ((Sim*)h)->seed(XF_SEED_REFERENCE, 0, draws); }
int64_t xo_num_particles(void* h) { return (int64_t)((Sim*)h)->px.size(); }
void xo_set_particles(void* h, int64_t n, const double* px, const double* py, const double* pz, const double* vx,
                      const double* vy, const double* vz) {
    Sim* s = (Sim*)h;
    s->px.assign(px, px + n); s->py.assign(py, py + n); s->pz.assign(pz, pz + n);
    s->vx.assign(vx, vx + n); s->vy.assign(vy, vy + n); s->vz.assign(vz, vz + n);
}
void xo_get_particles(void* h, double* px, double* py, double* pz, double* vx, double* vy, double* vz) {
    Sim* s = (Sim*)h;
    std::copy(s->px.begin(), s->px.end(), px); std::copy(s->py.begin(), s->py.end(), py);
    std::copy(s->pz.begin(), s->pz.end(), pz); std::copy(s->vx.begin(), s->vx.end(), vx);
    std::copy(s->vy.begin(), s->vy.end(), vy); std::copy(s->vz.begin(), s->vz.end(), vz);
}
static void* field_ptr(Sim* s, int which, size_t* bytes) {
    switch (which) {
        case XF_FIELD_U: *bytes = s->u.size() * 8; return s->u.data();
        case XF_FIELD_V: *bytes = s->v.size() * 8; return s->v.data();
        case XF_FIELD_W: *bytes = s->w.size() * 8; return s->w.data();
        case XF_FIELD_USAVED: *bytes = s->us.size() * 8; return s->us.data();
        case XF_FIELD_VSAVED: *bytes = s->vs.size() * 8; return s->vs.data();
        case XF_FIELD_WSAVED: *bytes = s->ws.size() * 8; return s->ws.data();
        case XF_FIELD_MATERIAL: *bytes = s->mat.size(); return s->mat.data();
        case XF_FIELD_PHI: *bytes = s->phi.size() * 8; return s->phi.data();
        case XF_FIELD_PRESSURE: *bytes = s->pressure.size() * 8; return s->pressure.data();
        case XF_FIELD_RHS: *bytes = s->rhs.size() * 8; return s->rhs.data();
        case XF_FIELD_MIC_INV: *bytes = s->micinv.size() * 8; return s->micinv.data();
    }
    *bytes = 0;
    return nullptr;
}
int xo_set_field(void* h, int which, const void* src, size_t bytes) {
    size_t b;
    void* p = field_ptr((Sim*)h, which, &b);
    if (!p || b != bytes) return -1;
    std::memcpy(p, src, b);
    return 0;
}
int xo_get_field(void* h, int which, void* dst, size_t bytes) {
    size_t b;
    void* p = field_ptr((Sim*)h, which, &b);
    if (!p || b != bytes) return -1;
    std::memcpy(dst, p, b);
    return 0;
}
int xo_stage(void* h, int stage, double dt) { return ((Sim*)h)->stage(stage, dt); }
void xo_add_fluid_source(void* h, const double* c, double wx, double wy, double wz, const double* v) {
    ((Sim*)h)->add_fluid_source(c, wx, wy, wz, v);
}
void xo_set_source_seed(void* h, uint64_t seed) { ((Sim*)h)->src_state = seed; }
// one _applyFluidSource call with the given 31-bit draws (3 per new particle)
int xo_apply_sources(void* h, double dt, const uint32_t* draws) { return ((Sim*)h)->apply_sources(dt, draws); }
int xo_time_step(void* h, double dt, int mesh) { return ((Sim*)h)->time_step(dt, mesh != 0); }
int xo_frame_step(void* h, double dt) { return ((Sim*)h)->frame_step(dt); }
void xo_last_solve(void* h, xf_solve_stats* st) { *st = ((Sim*)h)->last; }
int64_t xo_surface_cells(void* h, int32_t* cells, uint8_t* cube, int64_t cap) {
    Sim* s = (Sim*)h;
    int64_t n = (int64_t)s->surf.size();
    for (int64_t i = 0; i < n && i < cap; ++i) { cells[i] = s->surf[i]; cube[i] = s->cube[i]; }
    return n;
}
double xo_particle_radius(void* h) { return ((Sim*)h)->particle_radius(); }

// Projection u' = u - G A^-1 B u on explicit arrays (in place) and its adjoint
// g -> g + B^T A^-1 (-G^T g) (in place).  Return the PCG status.
int xo_project(int ni, int nj, int nk, const uint8_t* mat, double* u, double* v, double* w, double tol,
               int max_iter, double tau, double sigma) {
    const long N = (long)ni * nj * nk;
    std::vector<double> A(7 * N), b(N), p(N);
    assemble(ni, nj, nk, mat, u, v, w, A.data(), b.data());
    int it = 0;
    double res = 0;
    int st = pcg(ni, nj, nk, A.data(), b.data(), p.data(), tol, max_iter, tau, sigma, &it, &res, nullptr);
    apply_pressure(ni, nj, nk, mat, p.data(), u, v, w);
    return st;
}
int xo_project_adjoint(int ni, int nj, int nk, const uint8_t* mat, double* gu, double* gv, double* gw,
                       double tol, int max_iter, double tau, double sigma) {
    const long N = (long)ni * nj * nk;
    std::vector<double> A(7 * N), b(N), pbar(N), bbar(N);
    assemble(ni, nj, nk, mat, gu, gv, gw, A.data(), b.data());  // A only
    apply_pressure_adjoint(ni, nj, nk, mat, gu, gv, gw, pbar.data());
    int it = 0;
    double res = 0;
    int st = pcg(ni, nj, nk, A.data(), pbar.data(), bbar.data(), tol, max_iter, tau, sigma, &it, &res, nullptr);
    rhs_adjoint(ni, nj, nk, mat, bbar.data(), gu, gv, gw);
    return st;
}

// Particle transfers and their adjoints on explicit particle arrays (n x 3),
// using the handle's grid fields and parameters.
void xo_sample_points(void* h, int64_t n, const double* pos, double* out, int saved) {
    Sim* s = (Sim*)h;
    const double* F[3] = {saved ? s->us.data() : s->u.data(), saved ? s->vs.data() : s->v.data(),
                          saved ? s->ws.data() : s->w.data()};
    for (int64_t p = 0; p < n; ++p) s->get_velocity(pos + 3 * p, out + 3 * p, F[0], F[1], F[2]);
}
void xo_sample_adjoint(void* h, int64_t n, const double* pos, const double* gvel, double* gpos, double* gu,
                       double* gv, double* gw, int saved) {
    Sim* s = (Sim*)h;
    const double* F[3] = {saved ? s->us.data() : s->u.data(), saved ? s->vs.data() : s->v.data(),
                          saved ? s->ws.data() : s->w.data()};
    double* gF[3] = {gu, gv, gw};
    for (int64_t p = 0; p < n; ++p) s->sample_adjoint(pos + 3 * p, gvel + 3 * p, F, gF, gpos + 3 * p);
}
void xo_advect_points(void* h, int64_t n, const double* pos, double dt, double* out) {
    Sim* s = (Sim*)h;
    for (int64_t p = 0; p < n; ++p) s->advect_point(pos + 3 * p, dt, out + 3 * p);
}
// VJP of the position update with its collision branch (move_points): gnew on
// the new positions -> gpos (written), gu/gv/gw (added)
void xo_advect_adjoint(void* h, int64_t n, const double* pos, double dt, const double* gnew, double* gpos,
                       double* gu, double* gv, double* gw) {
    Sim* s = (Sim*)h;
    double* gF[3] = {gu, gv, gw};
    for (int64_t p = 0; p < n; ++p) s->move_adjoint_full(pos + 3 * p, dt, gnew + 3 * p, gF, gpos + 3 * p);
}
// VJP of the extrapolation stage on the handle's material (in place)
void xo_extrapolate_adjoint(void* h, double* gu, double* gv, double* gw) {
    Sim* s = (Sim*)h;
    const int layers = (int)(s->prm.cfl + 5);
    double* g[3] = {gu, gv, gw};
    for (int a = 0; a < 3; ++a) s->extrapolate_adjoint_axis(layers, a, g[a]);
}
// the whole position update of the advection stage (RK4 / RK2 target, then the
// collision branch) on the handle's fields and material
void xo_move_points(void* h, int64_t n, const double* pos, double dt, double* out) {
    Sim* s = (Sim*)h;
    for (int64_t p = 0; p < n; ++p) s->move_point(pos + 3 * p, dt, out + 3 * p);
}
void xo_p2g_points(void* h, int64_t n, const double* pos, const double* vel, double* u, double* v, double* w) {
    Sim* s = (Sim*)h;
    double* out[3] = {u, v, w};
    for (int axis = 0; axis < 3; ++axis) {
        std::vector<double> wt((size_t)(s->ni + (axis == 0)) * (s->nj + (axis == 1)) * (s->nk + (axis == 2)));
        s->p2g_points(axis, n, pos, vel, out[axis], wt.data());
    }
}
// P2G of one axis for given particles with the weight sums of every face
// (wt) beside the normalised values (out)
void xo_p2g_weights(void* h, int axis, int64_t n, const double* pos, const double* vel, double* out, double* wt) {
    ((Sim*)h)->p2g_points(axis, n, pos, vel, out, wt);
}
void xo_density_adjoint(void* h, int64_t n, const double* pos, const double* gphi, double* gpos) {
    ((Sim*)h)->density_adjoint(n, pos, gphi, gpos);
}
void xo_p2g_adjoint(void* h, int64_t n, const double* pos, const double* vel, const double* gu, const double* gv,
                    const double* gw, double* gvel, double* gpos) {
    Sim* s = (Sim*)h;
    const double* g[3] = {gu, gv, gw};
    s->p2g_adjoint(n, pos, vel, g, gvel, gpos);
}

// mesh of the current density field and surface list (stage DENSITY first)
int64_t xo_mesh(void* h, int smooth, int64_t* ntris) {
    Sim* s = (Sim*)h;
    s->surface();
    s->mesh(smooth != 0);
    *ntris = (int64_t)s->mtris.size() / 3;
    return (int64_t)s->mverts.size() / 3;
}
void xo_get_mesh(void* h, double* verts, int32_t* tris, double* normals) {
    Sim* s = (Sim*)h;
    if (verts) std::memcpy(verts, s->mverts.data(), s->mverts.size() * sizeof(double));
    if (tris) std::memcpy(tris, s->mtris.data(), s->mtris.size() * sizeof(int32_t));
    if (normals) std::memcpy(normals, s->mnormals.data(), s->mnormals.size() * sizeof(double));
}

}  // extern "C"
