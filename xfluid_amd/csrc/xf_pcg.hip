// xf_pcg.hip -- MIC(0)-preconditioned CG on the variational pressure matrix,
// matrix-free over the 7-point stencil codes (pcg_solver.h:85-310,
// sparse_matrix.h:266-276, blas_wrapper.h:12-51).
//
// The preconditioner reproduces the reference's lexicographic MIC(0): the
// factorisation and both triangular sweeps are hyperplane recurrences (every
// input of cell (i,j,k) lies on plane i+j+k-1, or i+j+k+1 for the backward
// sweep), so any schedule that respects the planes reproduces the serial
// arithmetic bit for bit.  Two schedules exist:
//   * one launch per hyperplane (simple, launch-bound);
//   * a pipelined persistent kernel over (j,k) line tiles (xf_sweep.hip).
// Dot products accumulate in FP64 with a fixed blocked order (deterministic,
// but not the reference's serial order: PCG iterates match to rounding).
#include <vector>
#include <cmath>
#include <functional>
#include <cstdio>
#include "xf_internal.h"
#include <climits>

#define RED_BLOCKS 1024

__device__ inline bool present(int cls) { return cls == 1 || cls == 2; }

// ---------------------------------------------------------------------------
// MIC(0) factorisation, pcg_solver.h:125-185 restated per cell (see the oracle
// for the reduction of the pattern search): pivot d_c starts at A_cc and takes
// one update from each present lower neighbour column k in ascending column
// order (c-nn, c-ni, c-1):  d -= L*L;  d -= (tau*L)*missing, where missing is
// the sum of the other present entries of column k in ascending row order.
__global__ void k_mic_plane(Geo g, int t, int kmin, const uint16_t* __restrict__ code, double* __restrict__ inv,
                            double tau, double sigma) {
    long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int j = (int)(tid % g.nj), k = kmin + (int)(tid / g.nj);
    int i = t - j - k;
    if (k >= g.nk || i < 0 || i >= g.ni) return;
    long c = cidx(g, i, j, k);
    uint16_t cc = code[c];
    double diag = code_diag(cc);
    if (diag == 0) { inv[c] = 0.0; return; }
    double d = diag;
    int cl = code_dir(cc, D_ZM);
    if (present(cl)) {  // column c-nn: other rows k+1, k+ni (both above c)
        long kc = c - g.nn;
        uint16_t ck = code[kc];
        double ik = inv[kc];
        double L = code_coef(cl) * ik;
        double miss = 0;
        int a = code_dir(ck, D_XP), b = code_dir(ck, D_YP);
        if (present(a)) miss += code_coef(a) * ik;
        if (present(b)) miss += code_coef(b) * ik;
        d -= L * L;
        d -= tau * L * miss;
    }
    cl = code_dir(cc, D_YM);
    if (present(cl)) {  // column c-ni: rows k+1 (above c), k+nn (below c)
        long kc = c - g.ni;
        uint16_t ck = code[kc];
        double ik = inv[kc];
        double L = code_coef(cl) * ik;
        double miss = 0;
        int a = code_dir(ck, D_XP), b = code_dir(ck, D_ZP);
        if (present(a)) miss += code_coef(a) * ik;
        if (present(b)) miss += code_coef(b) * ik;
        d -= L * L;
        d -= tau * L * miss;
    }
    cl = code_dir(cc, D_XM);
    if (present(cl)) {  // column c-1: rows k+ni, k+nn (both below c)
        long kc = c - 1;
        uint16_t ck = code[kc];
        double ik = inv[kc];
        double L = code_coef(cl) * ik;
        double miss = 0;
        int a = code_dir(ck, D_YP), b = code_dir(ck, D_ZP);
        if (present(a)) miss += code_coef(a) * ik;
        if (present(b)) miss += code_coef(b) * ik;
        d -= L * L;
        d -= tau * L * miss;
    }
    inv[c] = (d < sigma * diag) ? 1 / sqrt(diag) : 1 / sqrt(d);
}

static void plane_range(const Geo& g, int t, int* kmin, long* nthreads) {
    int k0 = t - (g.ni - 1) - (g.nj - 1);
    if (k0 < 0) k0 = 0;
    int k1 = t < g.nk - 1 ? t : g.nk - 1;
    *kmin = k0;
    *nthreads = (long)g.nj * (k1 - k0 + 1);
}

template <class T>
__global__ void k_fill(long n, T* __restrict__ d, T v);

// Active rows' bounding box: one block-level min/max per block, then atomics.
__global__ void k_active_box(Geo g, const uint16_t* __restrict__ code, int* box) {
    __shared__ int red[6][XF_BLOCK / 64];
    int v[6] = {INT_MAX, -1, INT_MAX, -1, INT_MAX, -1};
    auto visit = [&](long c, uint16_t cc) __attribute__((always_inline)) {
        bool act = code_diag(cc) != 0.0;
        for (int d = 0; d < 6; ++d) act = act || present(code_dir(cc, d));
        if (act) {  // (32-bit index arithmetic: N < 2^30)
            const unsigned cu = (unsigned)c, row = cu / (unsigned)g.ni;
            const int i = (int)(cu - row * (unsigned)g.ni), j = (int)(row % (unsigned)g.nj), k = (int)(row / (unsigned)g.nj);
            v[0] = i < v[0] ? i : v[0]; v[1] = i > v[1] ? i : v[1];
            v[2] = j < v[2] ? j : v[2]; v[3] = j > v[3] ? j : v[3];
            v[4] = k < v[4] ? k : v[4]; v[5] = k > v[5] ? k : v[5];
        }
    };
    // grid-stride: few blocks, so few atomics; a thread's codes are loaded U at a
    // time, unconditionally (one round trip to memory for all of them, not one
    // per cell; a cell past the end visits the last cell again, which changes
    // no bound)
    constexpr int U = 8;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long c0 = blockIdx.x * (long)blockDim.x + threadIdx.x; c0 < g.N; c0 += U * stride) {
        long cs[U];
        uint16_t ccs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            cs[u] = c0 + u * stride < g.N ? c0 + u * stride : g.N - 1;
            ccs[u] = code[cs[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) visit(cs[u], ccs[u]);
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int d = 0; d < 6; ++d) {
            const int w = __shfl_xor(v[d], o, 64);
            v[d] = (d & 1) ? (w > v[d] ? w : v[d]) : (w < v[d] ? w : v[d]);
        }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int d = 0; d < 6; ++d) red[d][wv] = v[d];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        int m = red[d][0];
        for (int q = 1; q < XF_BLOCK / 64; ++q) m = (d & 1) ? (red[d][q] > m ? red[d][q] : m) : (red[d][q] < m ? red[d][q] : m);
        // the bounds only tighten, so a block whose value does not improve on one
        // already seen skips its atomic (most blocks, once a few have landed)
        const int cur = __hip_atomic_load(box + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((d & 1) ? (m >= 0 && m > cur) : (m != INT_MAX && m < cur)) {
            if (d & 1) atomicMax(box + d, m);
            else atomicMin(box + d, m);
        }
    }
}
__global__ void k_box_init(int* box) {
    if (threadIdx.x < 6) box[threadIdx.x] = (threadIdx.x & 1) ? -1 : INT_MAX;
}

bool pcg_active_box(const Geo& g, const uint16_t* code, PcgWork& w, int box[6], hipStream_t st) {
    k_box_init<<<1, 64, 0, st>>>(w.box_dev);
    k_active_box<<<RED_BLOCKS, XF_BLOCK, 0, st>>>(g, code, w.box_dev);
    hipMemcpyAsync(w.box_host, w.box_dev, 6 * sizeof(int), hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    for (int d = 0; d < 6; ++d) box[d] = w.box_host[d];
    return box[1] >= box[0] && box[3] >= box[2] && box[5] >= box[4];
}

// box: when given, only the hyperplanes that meet the active rows' bounding
// box are factored and every other pivot is set to zero (their rows are empty).
void pcg_mic0(const Geo& g, const uint16_t* code, double* inv, double tau, double sigma, const int* box,
              hipStream_t st) {
    int t0 = 0, t1 = g.ni + g.nj + g.nk - 3;
    if (box) {
        k_fill<double><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(g.N, inv, 0.0);
        t0 = box[0] + box[2] + box[4];
        t1 = box[1] + box[3] + box[5];
    }
    for (int t = t0; t <= t1; ++t) {
        int kmin;
        long nt;
        plane_range(g, t, &kmin, &nt);
        k_mic_plane<<<xf_blocks(nt), XF_BLOCK, 0, st>>>(g, t, kmin, code, inv, tau, sigma);
    }
}

// ---------------------------------------------------------------------------
// solve_lower (pcg_solver.h:192-204) from cell c's side:
//   q_c = (((r_c - L_{c,c-nn} q_{c-nn}) - L_{c,c-ni} q_{c-ni}) - L_{c,c-1} q_{c-1}) * inv_c
// with L_{c,k} = A_{c,k} * inv_k.
template <class T>
__global__ void k_fwd_plane(Geo g, int t, int kmin, const uint16_t* __restrict__ code, const T* __restrict__ inv,
                            const T* __restrict__ r, T* __restrict__ q, const PcgScalars* S, int ax) {
    if (S && S->done) return;
    long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int j = (int)(tid % g.nj), k = kmin + (int)(tid / g.nj);
    int i = t - j - k;
    if (k >= g.nk || i < 0 || i >= g.ni) return;
    long c = cidx(g, i, j, k);
    uint16_t cc = code[c];
    T v = r[c];
    if constexpr (sizeof(T) == 4) {
        // FP32: the 8x8-tile sweep's order with lines along axis ax (the line's
        // previous cell, then the lanes' A and B neighbours: -i, -j, -k for ax 0,
        // -k, -i, -j for ax 2), fused multiply-adds
        const int d[3] = {ax == 2 ? D_ZM : D_XM, ax == 2 ? D_XM : D_YM, ax == 2 ? D_YM : D_ZM};
        const long o[3] = {ax == 2 ? g.nn : 1, ax == 2 ? 1 : (long)g.ni, ax == 2 ? (long)g.ni : g.nn};
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int cl = code_dir(cc, d[e]);
            if (present(cl)) v = __builtin_fmaf(-((T)code_coef(cl) * inv[c - o[e]]), q[c - o[e]], v);
        }
        q[c] = v * inv[c];
        return;
    }
    int cl = code_dir(cc, D_ZM);
    if (present(cl)) v = v - ((T)code_coef(cl) * inv[c - g.nn]) * q[c - g.nn];
    cl = code_dir(cc, D_YM);
    if (present(cl)) v = v - ((T)code_coef(cl) * inv[c - g.ni]) * q[c - g.ni];
    cl = code_dir(cc, D_XM);
    if (present(cl)) v = v - ((T)code_coef(cl) * inv[c - 1]) * q[c - 1];
    q[c] = v * inv[c];
}

// solve_lower_transpose_in_place (pcg_solver.h:207-220) from cell c's side:
//   z_c = (((q_c - L_{c+1,c} z_{c+1}) - L_{c+ni,c} z_{c+ni}) - L_{c+nn,c} z_{c+nn}) * inv_c
template <class T>
__global__ void k_bwd_plane(Geo g, int t, int kmin, const uint16_t* __restrict__ code, const T* __restrict__ inv,
                            const T* __restrict__ q, T* __restrict__ z, const PcgScalars* S, int ax) {
    if (S && S->done) return;
    long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    int j = (int)(tid % g.nj), k = kmin + (int)(tid / g.nj);
    int i = t - j - k;
    if (k >= g.nk || i < 0 || i >= g.ni) return;
    long c = cidx(g, i, j, k);
    uint16_t cc = code[c];
    T ic = inv[c];
    T v = q[c];
    if constexpr (sizeof(T) == 4) {
        // FP32: fused multiply-adds in the 8x8-tile sweep's order (+i, +j, +k for
        // lines along i, +k, +i, +j along k)
        const int d[3] = {ax == 2 ? D_ZP : D_XP, ax == 2 ? D_XP : D_YP, ax == 2 ? D_YP : D_ZP};
        const long o[3] = {ax == 2 ? g.nn : 1, ax == 2 ? 1 : (long)g.ni, ax == 2 ? (long)g.ni : g.nn};
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int cl = code_dir(cc, d[e]);
            if (present(cl)) v = __builtin_fmaf(-((T)code_coef(cl) * ic), z[c + o[e]], v);
        }
        z[c] = v * ic;
        return;
    }
    int cl = code_dir(cc, D_XP);
    if (present(cl)) v = v - ((T)code_coef(cl) * ic) * z[c + 1];
    cl = code_dir(cc, D_YP);
    if (present(cl)) v = v - ((T)code_coef(cl) * ic) * z[c + g.ni];
    cl = code_dir(cc, D_ZP);
    if (present(cl)) v = v - ((T)code_coef(cl) * ic) * z[c + g.nn];
    z[c] = v * ic;
}

// dot_partial: when given, the 8x8-tile schedule also writes the tiles'
// partials of dot(z, r) (returns true then; the caller runs w8_dot_tiles
// otherwise, which gives the same partials)
template <class T>
// chunk_end (optional): an event to record at the end of the application; when
// the pipelined sweep runs untimed it rides on the sweep's own dispatch (no
// marker packet, whose cache release idles the GPU for about 6 us) and
// *chunk_rec is set, else the caller records it.
static bool precond_t(const Geo& g, const uint16_t* code, const T* inv, const T* r, T* q, T* z, const PcgScalars* S,
                      int sweep_kernel, const SweepSkew<T>* sk, int* flags, unsigned long long* epoch,
                      SweepTimer* tm, hipStream_t st, double* dot_partial = nullptr, FinRes* fin = nullptr,
                      hipEvent_t chunk_end = nullptr, bool* chunk_rec = nullptr) {
    // a pending residual test runs first: inside the 8x8-tile schedule's fused
    // kernel, else as k_finish here
    if (fin && !(sweep_kernel == 9 && sk)) {
        launch_finish_res(*fin, st);
        fin->ran = true;
    }
    // pipelined 8x8-tile schedule (xf_sweep_wave.hip) on the records prepared by the caller
    if (sweep_kernel == 9 && sk) {
        hipEvent_t a = nullptr, b = nullptr;
        if (tm && tm->on && tm->k++ % SWEEP_TIMER_STRIDE == 0 && tm->used + 2 <= tm->cap) {
            a = tm->ev[tm->used];
            b = tm->ev[tm->used + 1];
        }
        const bool ride = !b && chunk_end;
        if (sweep_precond_w8<T>(g, *sk, r, z, S, flags, epoch, a, ride ? chunk_end : b, st, dot_partial, fin)) {
            if (a) tm->used += 2;
            if (ride && chunk_rec) *chunk_rec = true;
            return dot_partial != nullptr;
        }
    }
    if (fin && !fin->ran) {
        launch_finish_res(*fin, st);
        fin->ran = true;
    }
    const int planes = g.ni + g.nj + g.nk - 2;
    for (int t = 0; t < planes; ++t) {
        int kmin;
        long nt;
        plane_range(g, t, &kmin, &nt);
        k_fwd_plane<T><<<xf_blocks(nt), XF_BLOCK, 0, st>>>(g, t, kmin, code, inv, r, q, S, sk ? sk->ax : 0);
    }
    for (int t = planes - 1; t >= 0; --t) {
        int kmin;
        long nt;
        plane_range(g, t, &kmin, &nt);
        k_bwd_plane<T><<<xf_blocks(nt), XF_BLOCK, 0, st>>>(g, t, kmin, code, inv, q, z, S, sk ? sk->ax : 0);
    }
    return false;
}

void pcg_precond(const Geo& g, const uint16_t* code, const double* inv, const double* r, double* q, double* z,
                 int sweep_kernel, PcgWork& w, hipStream_t st) {
    if (!pcg_active_box(g, code, w, w.sk64.box, st)) {
        w.sk64.box[0] = w.sk64.box[2] = w.sk64.box[4] = 0;
        w.sk64.box[1] = g.ni - 1; w.sk64.box[3] = g.nj - 1; w.sk64.box[5] = g.nk - 1;
    }
    w.sk64.ax = w8_choose_axis(g, w.sk64.box);
    if (sweep_kernel == 9) sweep_prepare_w8<double>(g, code, inv, w.sk64, st);
    // (the pipelined sweeps write z at the box's tiles only)
    k_fill<double><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(g.N, z, 0.0);
    precond_t<double>(g, code, inv, r, q, z, nullptr, sweep_kernel, &w.sk64, w.sweep_flags, &w.epoch, nullptr, st);
}

// ---------------------------------------------------------------------------
// y = A x (sparse_matrix.h:266-276): present entries in column order, each
// added as a*x to a sum starting at 0.  Optional FP64 partial dot(x, y).
template <class T, class Get>
__device__ inline T spmv_cell_g(const Geo& g, uint16_t cc, long c, Get get);
template <class T>
__device__ inline T spmv_cell(const Geo& g, uint16_t cc, long c, const T* __restrict__ x) {
    return spmv_cell_g<T>(g, cc, c, [x](long e) __attribute__((always_inline)) { return x[e]; });
}
// The seven operand cells of row c in column order (-k, -j, -i, c, +i, +j, +k),
// indices clamped into the grid: every operand is read unconditionally, so the
// loads of several cells can be in flight together.
__device__ inline void spmv_cols(const Geo& g, long c, long e[7]) {
    e[0] = c >= g.nn ? c - g.nn : c;
    e[1] = c >= g.ni ? c - g.ni : c;
    e[2] = c >= 1 ? c - 1 : c;
    e[3] = c;
    e[4] = c + 1 < g.N ? c + 1 : c;
    e[5] = c + g.ni < g.N ? c + g.ni : c;
    e[6] = c + g.nn < g.N ? c + g.nn : c;
}
// row c of A times the operands x[0..6] (spmv_cols order): the present entries
// added in column order
template <class T>
__device__ inline T spmv_row(uint16_t cc, const T x[7]) {
    T y = 0;
    int cl = code_dir(cc, D_ZM);
    if (present(cl)) y += (T)code_coef(cl) * x[0];
    cl = code_dir(cc, D_YM);
    if (present(cl)) y += (T)code_coef(cl) * x[1];
    cl = code_dir(cc, D_XM);
    if (present(cl)) y += (T)code_coef(cl) * x[2];
    double dg = code_diag(cc);
    if (dg != 0.0) y += (T)dg * x[3];
    cl = code_dir(cc, D_XP);
    if (present(cl)) y += (T)code_coef(cl) * x[4];
    cl = code_dir(cc, D_YP);
    if (present(cl)) y += (T)code_coef(cl) * x[5];
    cl = code_dir(cc, D_ZP);
    if (present(cl)) y += (T)code_coef(cl) * x[6];
    return y;
}
// the same with the operand x[e] given by get(e)
template <class T, class Get>
__device__ inline T spmv_cell_g(const Geo& g, uint16_t cc, long c, Get get) {
    long e[7];
    spmv_cols(g, c, e);
    T x[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) x[i] = get(e[i]);
    return spmv_row<T>(cc, x);
}

template <class T>
__global__ void k_spmv_dot(Geo g, const uint16_t* __restrict__ code, const T* __restrict__ x, T* __restrict__ y,
                           double* partial, const PcgScalars* S, BoxRange bx) {
    __shared__ double sh[16];
    if (S && S->done) return;
    double acc = 0.0;
    const unsigned stride = gridDim.x * blockDim.x;
    // (each XCD walks a contiguous part of the box: neighbours' rows stay in its L2)
    unsigned t = (unsigned)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    BoxWalk w;
    bw_init(bx, t, w);
    // four cells of the grid-stride sequence at a time (the same cells, in the
    // same order, as one at a time)
    for (; t + 3 * stride < bx.n; t += 4 * stride) {
        long cs[4];
        uint16_t cc[4];
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            cs[u] = bw_cell(bx, w);
            bw_next(bx, w);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) cc[u] = code[cs[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = spmv_cell<T>(g, cc[u], cs[u], x);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            y[cs[u]] = v[u];
            acc += (double)x[cs[u]] * (double)v[u];
        }
    }
    for (; t < bx.n; t += stride) {
        const long c = bw_cell(bx, w);
        bw_next(bx, w);
        T v = spmv_cell<T>(g, code[c], c, x);
        y[c] = v;
        acc += (double)x[c] * (double)v;
    }
    acc = block_reduce<double, false>(acc, sh);
    if (partial && threadIdx.x == 0) partial[blockIdx.x] = acc;
}

static BoxRange full_range(const Geo& g) {
    const int all[6] = {0, g.ni - 1, 0, g.nj - 1, 0, g.nk - 1};
    return box_range(g.ni, g.nj, g.nk, all, RED_BLOCKS * XF_BLOCK);
}

void pcg_spmv(const Geo& g, const uint16_t* code, const double* x, double* y, hipStream_t st) {
    k_spmv_dot<double><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(g, code, x, y, nullptr, nullptr, full_range(g));
}

// r = b - A x (FP64), partial max |r| -- the refinement residual.
// guard (optional): runs only once that solve has converged (enqueued ahead of
// knowing it, see pcg_core's `post`)
__global__ void k_residual(Geo g, const uint16_t* __restrict__ code, const double* __restrict__ b,
                           const double* __restrict__ x, double* __restrict__ r, double* partial,
                           const PcgScalars* guard = nullptr) {
    __shared__ double sh[16];
    if (guard && !guard->done) return;
    double m = 0.0;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < g.N; c += (long)gridDim.x * blockDim.x) {
        double v = b[c] - spmv_cell<double>(g, code[c], c, x);
        r[c] = v;
        double a = fabs(v);
        m = a > m ? a : m;
    }
    m = block_reduce<double, true>(m, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

template <bool MAX>
__device__ inline double reduce_partials(const double* p, int n, double* sh) {
    // a thread's first four partials are loaded together (they were just written by
    // another kernel: each load is a round trip to memory), then added in index order
    constexpr int U = 4;
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = threadIdx.x + u * blockDim.x;
        a[u] = i < n ? p[i] : 0.0;
    }
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (threadIdx.x + u * blockDim.x < n) v = MAX ? (a[u] > v ? a[u] : v) : v + a[u];
    for (int i = threadIdx.x + U * blockDim.x; i < n; i += blockDim.x) v = MAX ? (p[i] > v ? p[i] : v) : v + p[i];
    return block_reduce<double, MAX>(v, sh);
}

// x += alpha s; r += (-alpha) z; partial max |r|   (blas_wrapper.h:41-51),
// z = A s here.  The alpha step (alpha = rho / dot(s, A s), pcg_solver.h:262)
// runs in every block from the spmv's partials (pin): each block reduces them
// in the same order, so all agree, and block 0 records sz and alpha.
template <class T>
__global__ void k_axpy_max(BoxRange bx, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ s,
                           const T* __restrict__ z, const double* __restrict__ pin, int npin, double* partial,
                           PcgScalars* S) {
    __shared__ double sh[16];
    __shared__ double alpha_sh;
    // the solver state is read with the operands below (not ahead of them, one
    // round trip each): a finished solve returns after their requests, before
    // any store
    const int done = S->done;
    const double rho_next = S->rho_next;
    const unsigned stride = gridDim.x * blockDim.x;
    // (each XCD walks a contiguous part of the box: neighbours' rows stay in its L2)
    unsigned t = (unsigned)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    BoxWalk w;
    bw_init(bx, t, w);
    // the operands of this thread's first PF cells are requested before the
    // spmv's partials, so the two round trips to memory overlap (an absent
    // cell reads cell 0 and is not written)
    constexpr int PF = 2;
    long c0[PF];
    T x0[PF], r0[PF], s0[PF], z0[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        c0[u] = t + u * stride < bx.n ? bw_cell(bx, w) : 0;
        bw_next(bx, w);
        x0[u] = x[c0[u]]; r0[u] = r[c0[u]]; s0[u] = s[c0[u]]; z0[u] = z[c0[u]];
    }
    if (done) return;
    const double sz = reduce_partials<false>(pin, npin, sh);
    if (threadIdx.x == 0) {
        alpha_sh = rho_next / sz;
        // (rho_next is this iteration's rho; nothing reads rho while this runs)
        if (blockIdx.x == 0) { S->sz = sz; S->alpha = alpha_sh; S->rho = rho_next; }
    }
    __syncthreads();
    const double alpha = alpha_sh;
    const T a = (T)alpha, na = (T)(-alpha);
    double m = 0.0;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        if (t + u * stride < bx.n) {
            x[c0[u]] = x0[u] + a * s0[u];
            const T rv = r0[u] + na * z0[u];
            r[c0[u]] = rv;
            const double av = fabs((double)rv);
            m = av > m ? av : m;
        }
    }
    for (t += PF * stride; t < bx.n; t += stride) {
        const long c = bw_cell(bx, w);
        bw_next(bx, w);
        x[c] += a * s[c];
        T rv = r[c] + na * z[c];
        r[c] = rv;
        double av = fabs((double)rv);
        m = av > m ? av : m;
    }
    m = block_reduce<double, true>(m, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

template <class T>
__global__ void k_dot(BoxRange bx, const T* __restrict__ a, const T* __restrict__ b, double* partial,
                      const PcgScalars* S) {
    __shared__ double sh[16];
    if (S && S->done) return;
    double acc = 0.0;
    const unsigned stride = gridDim.x * blockDim.x;
    // (each XCD walks a contiguous part of the box: neighbours' rows stay in its L2)
    unsigned t = (unsigned)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    BoxWalk w;
    bw_init(bx, t, w);
    for (; t < bx.n; t += stride) {
        const long c = bw_cell(bx, w);
        bw_next(bx, w);
        acc += (double)a[c] * (double)b[c];
    }
    acc = block_reduce<double, false>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Start of a PCG solve: r = (T) rsrc, x = 0, z = 0 and the partials of max |r|
// (the copy, fills and reduction of separate kernels, in one pass).
template <class T, class Src>
__global__ void k_pcg_start(long n, T* __restrict__ r, const Src* __restrict__ rsrc, T* __restrict__ x,
                            T* __restrict__ z, double* partial) {
    __shared__ double sh[16];
    double m = 0.0;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
        const T v = (T)rsrc[c];
        r[c] = v;
        x[c] = (T)0;
        z[c] = (T)0;
        const double a = fabs((double)v);
        m = a > m ? a : m;
    }
    m = block_reduce<double, true>(m, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

template <class T>
__global__ void k_absmax(long n, const T* __restrict__ a, double* partial) {
    __shared__ double sh[16];
    double m = 0.0;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
        double v = fabs((double)a[c]);
        m = v > m ? v : m;
    }
    m = block_reduce<double, true>(m, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

// s_new = z + beta s_old (pcg_solver.h:279-281; the reference computes
// z_i + beta*s_i, then swaps) fused with the next
// iteration's y = A s_new and the partials of dot(s_new, y): each cell's
// s_new (and its neighbours') is formed with the T arithmetic of the separate
// update, so the values are the same.  The beta step (beta = rho_new / rho)
// runs in every block from the dot(z, r) partials (pin); block 0 records it
// and leaves the new rho in rho_next (other blocks still read rho here).
template <class T>
__global__ void k_spmv_upd(Geo g, const uint16_t* __restrict__ code, const T* __restrict__ z,
                           const T* __restrict__ s_old, T* __restrict__ s_new, T* __restrict__ y,
                           const double* __restrict__ pin, int npin, double* partial, PcgScalars* S, BoxRange bx) {
    __shared__ double sh[16];
    __shared__ double beta_sh;
    // the solver state is read with the operands below, and a finished solve
    // returns after their requests, before any store (as in k_axpy_max)
    const int was = S->done;
    const double rho = S->rho;
    const unsigned stride = gridDim.x * blockDim.x;
    // (each XCD walks a contiguous part of the box: neighbours' rows stay in its L2)
    unsigned t = (unsigned)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    BoxWalk w;
    bw_init(bx, t, w);
    // the codes and operands (z, s_old of the seven columns) of this thread's
    // first PF cells are requested before the dot(z, r) partials, so the two
    // round trips to memory overlap (an absent cell reads cell 0 and is not
    // written)
    constexpr int PF = 2;
    long c0[PF];
    uint16_t cc0[PF];
    T z0[PF][7], o0[PF][7];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        c0[u] = t + u * stride < bx.n ? bw_cell(bx, w) : 0;
        bw_next(bx, w);
        cc0[u] = code[c0[u]];
        long e[7];
        spmv_cols(g, c0[u], e);
#pragma unroll
        for (int i = 0; i < 7; ++i) { z0[u][i] = z[e[i]]; o0[u][i] = s_old[e[i]]; }
    }
    if (was) return;
    const double v = reduce_partials<false>(pin, npin, sh);
    if (threadIdx.x == 0) {
        beta_sh = v / rho;
        if (blockIdx.x == 0) { S->rho_new = v; S->beta = beta_sh; S->rho_next = v; }
    }
    __syncthreads();
    const T beta = (T)beta_sh;
    auto sv = [=](long e) __attribute__((always_inline)) { return z[e] + beta * s_old[e]; };
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        if (t + u * stride < bx.n) {
            T xs[7];
#pragma unroll
            for (int i = 0; i < 7; ++i) xs[i] = z0[u][i] + beta * o0[u][i];
            const T yv = spmv_row<T>(cc0[u], xs);
            s_new[c0[u]] = xs[3];
            y[c0[u]] = yv;
            acc += (double)xs[3] * (double)yv;
        }
    }
    for (t += PF * stride; t + 3 * stride < bx.n; t += 4 * stride) {
        long cs[4];
        uint16_t cc[4];
        T yv[4], sc[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            cs[u] = bw_cell(bx, w);
            bw_next(bx, w);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) cc[u] = code[cs[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            yv[u] = spmv_cell_g<T>(g, cc[u], cs[u], sv);
            sc[u] = sv(cs[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            s_new[cs[u]] = sc[u];
            y[cs[u]] = yv[u];
            acc += (double)sc[u] * (double)yv[u];
        }
    }
    for (; t < bx.n; t += stride) {
        const long c = bw_cell(bx, w);
        bw_next(bx, w);
        const T yv = spmv_cell_g<T>(g, code[c], c, sv), sc = sv(c);
        s_new[c] = sc;
        y[c] = yv;
        acc += (double)sc * (double)yv;
    }
    acc = block_reduce<double, false>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

template <class D, class Sx>
__global__ void k_copy(long n, D* __restrict__ d, const Sx* __restrict__ s, const PcgScalars* S) {
    if (S && S->done) return;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x)
        d[c] = (D)s[c];
}
template <class T>
__global__ void k_fill(long n, T* __restrict__ d, T v) {
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) d[c] = v;
}
__global__ void k_add_f2d(long n, double* __restrict__ x, const float* __restrict__ c,
                          const PcgScalars* guard = nullptr) {
    if (guard && !guard->done) return;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] += (double)c[i];
}

// ---- final reductions (one block, fixed order) and scalar updates ----

enum { FIN_RES0, FIN_RHO0, FIN_ALPHA, FIN_RES, FIN_BETA };
// mirror (optional): pinned host copy of S, written at the end (the host reads
// it after synchronising the stream, with no copy kernel in between)
// eflag_out (optional, pinned): receives *eflag, the sweeps' error word, with the mirror
__global__ void k_finish(int what, const double* partial, int n, PcgScalars* S, double tol_factor, int it,
                         PcgScalars* mirror = nullptr, const int* eflag = nullptr, int* eflag_out = nullptr,
                         const PcgScalars* guard = nullptr) {
    __shared__ double sh[16];
    if (guard && !guard->done) return;
    if (what != FIN_RES0 && S->done) {
        if (mirror && threadIdx.x == 0) *mirror = *S;
        if (eflag_out && threadIdx.x == 0) *eflag_out = *eflag;
        return;
    }
    const bool mx = (what == FIN_RES0 || what == FIN_RES);
    double v = mx ? reduce_partials<true>(partial, n, sh) : reduce_partials<false>(partial, n, sh);
    if (threadIdx.x != 0) return;
    switch (what) {
        case FIN_RES0:
            S->done = 0; S->iters = 0; S->status = 0; S->it = 0;
            S->res0 = v; S->res = v;
            S->tol = tol_factor * v;
            if (v == 0) S->done = 1;
            break;
        case FIN_RHO0:
            S->rho = v;
            S->rho_next = v;
            if (v == 0 || v != v) { S->done = 1; S->status = 2; S->iters = 0; }
            break;
        case FIN_ALPHA:
            S->sz = v;
            S->alpha = S->rho / v;
            break;
        case FIN_RES:
            S->res = v;
            S->it = it + 1;
            if (v <= S->tol) { S->done = 1; S->iters = it + 1; }
            break;
        case FIN_BETA:
            S->rho_new = v;
            S->beta = v / S->rho;
            S->rho = v;
            break;
    }
    if (mirror) *mirror = *S;
    if (eflag_out) *eflag_out = *eflag;
}

void launch_finish_res(const FinRes& f, hipStream_t st) {
    k_finish<<<1, 1024, 0, st>>>(FIN_RES, f.partial, f.n, f.S, 0.0, f.it, f.mirror);
}

// ---------------------------------------------------------------------------
// One PCG solve (pcg_solver.h:243-285) of A x = rhs with T-precision vectors.
// x starts at zero and r at rhs (rounded to T), both set here.
// Returns the solver state after the converged (or last) iteration.  A
// converged solve returns as soon as the host has seen its status: the
// iterations enqueued after it, which return at once, may still be in flight
// on the stream, ahead of whatever the caller enqueues next.
// post (optional): enqueues the caller's work after the solve, each kernel of it
// guarded by S->done (it runs only if the solve has converged by then).  It is
// called once, right after the chunk that is predicted to converge, so that
// this work follows the solve with no wait for the host; *post_ran tells the
// caller whether it was enqueued behind the converged chunk (else the caller
// enqueues it itself).
template <class T, class Src>
static PcgScalars pcg_core(const Geo& g, const uint16_t* code, const T* inv, const Src* rhs, T* x, T* r, T* s, T* s2,
                           T* z, T* q,
                     double* partial, int half, PcgScalars* S, PcgScalars* S_host, hipEvent_t* chunk_ev,
                     double tol_factor,
                     int max_iter, int sweep_kernel, const SweepSkew<T>* sk, int* flags, unsigned long long* epoch,
                     SweepTimer* tm, hipStream_t st, const std::function<void()>* post = nullptr,
                     bool* post_ran = nullptr) {
    if (post_ran) *post_ran = false;
    const long N = g.N;
    const int tm_start = tm ? tm->used : 0;
    if (tm) tm->k = 0;
    double* partial2 = partial + half;
    // the iteration kernels walk only the rows of the active rows' bounding box
    // (every vector is zero elsewhere: b has no rows there and x starts at zero;
    // z is written at the box's tiles only, and never read outside the box)
    const BoxRange bxr = sk ? box_range(g.ni, g.nj, g.nk, sk->box, RED_BLOCKS * XF_BLOCK) : full_range(g);
    // dot(z, r) comes as the partials of the box's 8x8 tiles (ndot of them) from the
    // pipelined backward sweep or w8_dot_tiles, whichever schedule ran: the same bits
    const int ndot = sk ? w8_box_tiles(g, sk->ax, sk->box) : RED_BLOCKS;
    auto dot_zr = [&]() {
        if (sk) w8_dot_tiles<T>(g, sk->ax, sk->box, z, r, partial, S, st);
        else k_dot<T><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(bxr, z, r, partial, S);
    };
    k_pcg_start<T, Src><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(N, r, rhs, x, z, partial);
    k_finish<<<1, 1024, 0, st>>>(FIN_RES0, partial, RED_BLOCKS, S, tol_factor, 0);
    if (!precond_t<T>(g, code, inv, r, q, z, S, sweep_kernel, sk, flags, epoch, tm, st, partial)) dot_zr();
    k_finish<<<1, 1024, 0, st>>>(FIN_RHO0, partial, ndot, S, tol_factor, 0);
    k_copy<T, T><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(N, s, z, S);
    const int CH = 4;
    int it = 0;
    // s alternates between two buffers: iteration it >= 1 forms its s from z and
    // the previous s inside the spmv (k_spmv_upd), A s goes to q
    T* sc = s;
    T* sn = s2;
    // Chunks of CH iterations; the host enqueues chunk c + 1 before it waits for
    // chunk c (whose last residual step writes its status into the pinned slot
    // S_host[c & 1 ? 2 : 0]), so the GPU does not idle while the host checks.
    // After convergence the kernels of the chunk in flight return at once.
    auto launch_chunk = [&](int c, int n) {
        int end = it + n < max_iter ? it + n : max_iter;
        PcgScalars* slot = S_host + ((c & 1) ? 2 : 0);
        bool rec = false;  // the chunk's event rode on its last sweep
        for (; it < end; ++it) {
            // (the spmv and dot partials go to `partial`, the residual's to `partial2`)
            if (it == 0) {
                k_spmv_dot<T><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(g, code, sc, q, partial, S, bxr);
            } else {
                k_spmv_upd<T><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(g, code, z, sc, sn, q, partial, ndot, partial2, S,
                                                               bxr);
                T* t = sc;
                sc = sn;
                sn = t;
            }
            k_axpy_max<T><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(bxr, x, r, sc, q, it == 0 ? partial : partial2, RED_BLOCKS,
                                                            it == 0 ? partial2 : partial, S);
            FinRes fin{it == 0 ? partial2 : partial, RED_BLOCKS, it, S, it == end - 1 ? slot : nullptr, false};
            if (!precond_t<T>(g, code, inv, r, q, z, S, sweep_kernel, sk, flags, epoch, tm, st, partial, &fin,
                              it == end - 1 ? chunk_ev[c & 1] : nullptr, &rec))
                dot_zr();
        }
        if (!rec) hipEventRecord(chunk_ev[c & 1], st);
    };
    // Chunk sizes: CH iterations, fewer once the decay of max |r| between the
    // last two chunks predicts convergence sooner (one iteration of margin), so
    // that few iterations are enqueued past it (each one's three launches
    // return at once but still cost about 4 us each).  When the chunks enqueued
    // so far are predicted to reach it, no further chunk is enqueued until the
    // last one's status is in; if it has not converged then, the GPU idles for
    // the host's turnaround while the next chunk is enqueued.
    int launched = 0, want = max_iter;
    int end_it[2] = {0, 0};  // iteration count at the end of chunk c, at [c & 1]
    auto enqueue = [&](int n) {
        const int c = launched++;
        launch_chunk(c, n);
        end_it[c & 1] = it;
        want -= n;
    };
    enqueue(CH);
    int chunk = 0;
    double prev_res = -1.0;
    int prev_it = 0;
    // the host reads each chunk's slot into its own copy: the chunk enqueued
    // next may already be writing the other slot
    PcgScalars hs{};
    int post_at = 0;  // chunks enqueued ahead of the caller's work (0: not enqueued)
    for (;;) {
        // at most two chunks in flight: this one and the next
        if (launched == chunk + 1 && it < max_iter && want > 0) enqueue(want < CH ? want : CH);
        if (post && !post_at && (want <= 0 || it >= max_iter)) {
            (*post)();
            post_at = launched;
        }
        hipEventSynchronize(chunk_ev[chunk & 1]);
        hs = S_host[(chunk & 1) ? 2 : 0];
        if (hs.done && post_ran) *post_ran = chunk < post_at;
        if (hs.done || (it >= max_iter && launched == chunk + 1)) break;
        // iterations still needed after this chunk's test, from the decay rate of
        // max |r| since the previous chunk's (all of max_iter while unknown)
        int need = max_iter;
        if (prev_res > 0.0 && hs.res > 0.0 && hs.res < prev_res && hs.it > prev_it && hs.tol > 0.0) {
            const double rate = std::log(hs.res / prev_res) / (double)(hs.it - prev_it);
            const double k = std::log(hs.tol / hs.res) / rate;
            need = k < (double)max_iter ? (int)std::ceil(k) + 1 : max_iter;
        }
        prev_res = hs.res;
        prev_it = hs.it;
        // (iterations already enqueued after this chunk)
        const int ahead = launched > chunk + 1 ? end_it[(chunk + 1) & 1] - end_it[chunk & 1] : 0;
        want = need - ahead;
        ++chunk;
        // nothing in flight (the prediction fell short): enqueue at least one iteration
        if (launched == chunk && it < max_iter) enqueue(want > 1 ? (want < CH ? want : CH) : 1);
    }
    if (!hs.done) {
        // (the stream is idle after this wait: no kernel writes the slots then)
        hipStreamSynchronize(st);
        hipMemcpyAsync(S_host, S, sizeof(PcgScalars), hipMemcpyDeviceToHost, st);
        hipStreamSynchronize(st);
        hs = *S_host;
        if (!hs.done) { hs.status = 1; hs.iters = max_iter; }
    }
    if (tm && tm->used > tm_start) {
        // keep the timings of the applications that ran: the initial one and one per
        // iteration before the converged one (later launches returned at once);
        // applications 0, stride, 2 stride, ... were timed
        const int apps = hs.res0 == 0 ? 0 : hs.status == 1 ? 1 + max_iter : hs.iters;
        const int timed = apps > 0 ? (apps - 1) / SWEEP_TIMER_STRIDE + 1 : 0;
        if (tm->used > tm_start + 2 * timed) tm->used = tm_start + 2 * timed;
    }
    return hs;
}

PcgResult pcg_solve(const Geo& g, const uint16_t* code, PcgWork& w, double tol_factor, int max_iter, double tau,
                    double sigma, int precision, double inner_tol, int max_outer, int sweep_kernel, hipStream_t st) {
    const long N = g.N;
    if (tol_factor < 1e-30) tol_factor = 1e-30;
    PcgResult res{};
    int box[6];
    if (!pcg_active_box(g, code, w, box, st)) {
        box[0] = box[2] = box[4] = 0;
        box[1] = g.ni - 1; box[3] = g.nj - 1; box[5] = g.nk - 1;
    }
    for (int d = 0; d < 6; ++d) w.sk64.box[d] = w.sk32.box[d] = box[d];
    w.sk64.ax = w.sk32.ax = w8_choose_axis(g, box);
    res.box_cells = (long)(box[1] - box[0] + 1) * (box[3] - box[2] + 1) * (box[5] - box[4] + 1);
    if (!(sweep_kernel == 9 && mic0_w8(g, code, w.inv, tau, sigma, w.sk64, w.sweep_flags, &w.epoch, st)))
        pcg_mic0(g, code, w.inv, tau, sigma, box, st);
    if (precision == 0) {
        if (sweep_kernel == 9) sweep_prepare_w8<double>(g, code, w.inv, w.sk64, st);
        const PcgScalars hs = pcg_core<double, double>(g, code, w.inv, w.b, w.x, w.r, w.s, w.s2, w.z, w.q, w.partial,
                                                       w.npartial, w.S, w.S_host, w.chunk_ev, tol_factor,
                         max_iter,
                         sweep_kernel, &w.sk64, w.sweep_flags, &w.epoch, &w.timer, st);
        res.status = hs.status;
        res.iterations = hs.iters;
        res.outer = 1;
        res.residual = hs.res;
        res.rhs_norm = hs.res0;
        return res;
    }
    // Mixed precision: FP32 inner PCG (SpMV, sweeps, axpy in FP32; dots in
    // FP64; pivots factored in FP64 then rounded) inside FP64 iterative
    // refinement on the true residual r = b - A x.
    PcgScalars* S_outer = w.S + 1;
    PcgScalars* S_outer_h = w.S_host + 1;
    k_copy<float, double><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(N, w.inv32, w.inv, nullptr);
    if (sweep_kernel == 9) sweep_prepare_w8<float>(g, code, w.inv32, w.sk32, st);
    // x = 0, r = b, max |r| (w.z, unused by the mixed solve, is cleared on the way)
    k_pcg_start<double, double><<<RED_BLOCKS, XF_BLOCK, 0, st>>>(N, w.r, w.b, w.x, w.z, w.partial);
    // (no wait here: the host learns max |b| and the tolerance from the mirror of
    // the first refinement step below, which copies all of S_outer)
    k_finish<<<1, 1024, 0, st>>>(FIN_RES0, w.partial, RED_BLOCKS, S_outer, tol_factor, 0);
    double tol = 0.0;
    int total = 0, outer = 0;
    res.status = 1;
    while (outer < max_outer && total < max_iter * max_outer) {
        // The inner solve reduces its residual by inner_tol, or only by what is
        // still missing to the outer tolerance (with a factor 2 of margin for the
        // FP32 recurrence drifting from the true residual), whichever is less;
        // for the first one, what is missing is tol_factor itself.
        const double need = outer == 0 ? 0.5 * tol_factor : 0.5 * tol / res.residual;
        const double itol = need > inner_tol ? (need < 0.5 ? need : 0.5) : inner_tol;
        // the refinement step: x += x32, r = b - A x and its max |r| (with the
        // solve's error word: every sweep that ran did so before it; the
        // iterations enqueued after convergence return at once).  Guarded by the
        // inner solve's status, it is enqueued behind the chunk predicted to
        // converge, so that it follows the solve without a wait for the host.
        auto refine = [&](const PcgScalars* guard) {
            k_add_f2d<<<RED_BLOCKS, XF_BLOCK, 0, st>>>(N, w.x, w.x32, guard);
            k_residual<<<RED_BLOCKS, XF_BLOCK, 0, st>>>(g, code, w.b, w.x, w.r, w.partial, guard);
            k_finish<<<1, 1024, 0, st>>>(FIN_RES, w.partial, RED_BLOCKS, S_outer, 0.0, 0, S_outer_h, w.sweep_flags,
                                         w.err_host, guard);
        };
        const std::function<void()> post = [&]() { refine(w.S); };
        bool post_ran = false;
        const PcgScalars hs = pcg_core<float, double>(g, code, w.inv32, w.r, w.x32, w.r32, w.s32, w.s2_32, w.z32,
                                                      w.q32, w.partial, w.npartial, w.S, w.S_host,
                        w.chunk_ev, itol,
                        max_iter, sweep_kernel, &w.sk32, w.sweep_flags, &w.epoch, &w.timer, st, &post, &post_ran);
        total += hs.iters;
        ++outer;
        if (!post_ran) refine(nullptr);
        hipStreamSynchronize(st);
        res.sweep_err = *(volatile int*)w.err_host;
        res.err_known = true;
        res.rhs_norm = S_outer_h->res0;
        tol = S_outer_h->tol;
        res.residual = S_outer_h->res;
        if (S_outer_h->res <= tol) { res.status = 0; break; }
        if (hs.status == 2) { res.status = 2; break; }
    }
    res.iterations = total;
    res.outer = outer;
    return res;
}
