// xf_sweep.hip -- pipelined MIC(0) triangular sweeps (persistent line tiles).
//
// The forward sweep q_c = (((r - L q_{c-nn}) - L q_{c-ni}) - L q_{c-1}) inv_c
// (pcg_solver.h:192-204) and the backward sweep (pcg_solver.h:207-220) are
// hyperplane recurrences.  Schedule: the (j,k) plane is cut into TJ x TK tiles
// of i-lines; one workgroup owns one tile, one thread one line, and walks i
// with a skew of jl+kl steps, so at step s thread (jl,kl) computes the cell of
// plane s + (tile origin): its -j / -k neighbours were computed one step
// earlier by the neighbouring threads (exchanged through LDS), its -i
// neighbour by itself.  Tiles form a 2-D pipeline: a tile reads the edge
// lines of its -j and -k neighbour tiles from global memory after those
// publish enough progress.  Each cell is evaluated with exactly the
// operations of the serial sweep, so results are bit-identical to the
// per-hyperplane launches and to the reference.
//
// Progress counters carry an epoch in their high bits (no reset between
// sweeps).  Waits are bounded: on timeout the kernel records an error and
// finishes, and the host redoes the sweep with per-hyperplane launches.
#include <atomic>
#include "xf_internal.h"

#define PT_J 16
#define PT_K 16
#define PT_T (PT_J * PT_K)
#define PT_C 8                    // chunk: loads are issued one chunk ahead
#define PT_SPIN_LIMIT (1 << 22)  // polls (~0.5 s) before a wait is declared stuck

__device__ inline bool present_cls(int cls) { return cls == 1 || cls == 2; }

__device__ inline void tile_of_block(int b, int ntj, int ntk, bool fwd, int* J, int* K) {
    // enumerate tiles by diagonal J+K (upstream tiles first)
    int d = 0;
    for (;; ++d) {
        int j0 = d - (ntk - 1) > 0 ? d - (ntk - 1) : 0;
        int j1 = d < ntj - 1 ? d : ntj - 1;
        int cnt = j1 - j0 + 1;
        if (b < cnt) { *J = j0 + b; *K = d - *J; break; }
        b -= cnt;
    }
    if (!fwd) { *J = ntj - 1 - *J; *K = ntk - 1 - *K; }
}

template <class T, bool FWD>
__global__ __launch_bounds__(PT_T) void k_sweep_pipe(Geo g, int ntj, int ntk, const uint16_t* __restrict__ code,
                                                     const T* __restrict__ inv, const T* __restrict__ in,
                                                     T* __restrict__ out, const PcgScalars* S,
                                                     unsigned long long* prog, unsigned long long epoch, int* err) {
    __shared__ T lq[2][PT_K][PT_J];
    __shared__ T li[2][PT_K][PT_J];
    if (S && S->done) return;
    int J, K;
    tile_of_block(blockIdx.x, ntj, ntk, FWD, &J, &K);
    const int tid = threadIdx.x, jl = tid % PT_J, kl = tid / PT_J;
    const int j = J * PT_J + jl, k = K * PT_K + kl;
    const bool valid = j < g.nj && k < g.nk;
    // skew in steps: distance from the tile corner that starts first
    const int jd = FWD ? jl : (PT_J - 1 - jl), kd = FWD ? kl : (PT_K - 1 - kl);
    const int skew = jd + kd;
    const int total = g.ni + (PT_J - 1) + (PT_K - 1);
    // neighbour tiles this tile depends on
    const int uJ = FWD ? J - 1 : J + 1, uK = FWD ? K - 1 : K + 1;
    const bool hasJ = uJ >= 0 && uJ < ntj, hasK = uK >= 0 && uK < ntk;
    const int myTile = J * ntk + K;
    const unsigned long long base = epoch << 20;
    // edge lanes read the neighbour tile's edge line from global memory
    const bool edgeJ = (jd == 0) && hasJ, edgeK = (kd == 0) && hasK;
    const long line = valid ? (long)g.ni * j + g.nn * k : 0;
    const long offJ = FWD ? -(long)g.ni : (long)g.ni;
    const long offK = FWD ? -g.nn : g.nn;
    const int dJ = FWD ? D_YM : D_YP, dK = FWD ? D_ZM : D_ZP, dI = FWD ? D_XM : D_XP;

    // chunk registers: own r/inv/code and the edge-neighbour q/inv for the
    // PT_C steps of the current chunk; the next chunk is loaded one ahead
    T cr[PT_C], ci[PT_C], nr[PT_C], ni_[PT_C];
    uint16_t cc[PT_C], nc[PT_C];
    T ejq[PT_C], eji[PT_C], ekq[PT_C], eki[PT_C];
    T njq[PT_C], nji[PT_C], nkq[PT_C], nki[PT_C];
    __shared__ int knownJ, knownK;

    auto cell_i = [&](int s) { int il = s - skew; return FWD ? il : g.ni - 1 - il; };
    auto in_range = [&](int s) { int il = s - skew; return valid && il >= 0 && il < g.ni; };
    auto load_own = [&](int s0, T* r_, T* i_, uint16_t* c_) {
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            int s = s0 + u;
            if (in_range(s)) {
                long c = line + cell_i(s);
                r_[u] = in[c]; i_[u] = inv[c]; c_[u] = code[c];
            } else { r_[u] = 0; i_[u] = 0; c_[u] = 0; }
        }
    };
    auto load_edge = [&](int s0, bool on, long off, T* q_, T* i_) {
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            int s = s0 + u;
            if (on && in_range(s)) {
                long c = line + cell_i(s) + off;
                q_[u] = out[c]; i_[u] = inv[c];
            } else { q_[u] = 0; i_[u] = 0; }
        }
    };
    // wait until neighbour tile t has completed `need` steps
    auto wait_for = [&](int t, int need, int* known) {
        if (*known >= need) return;
        unsigned long long want = base + (unsigned long long)need, v;
        int spins = 0;
        while ((v = __hip_atomic_load(&prog[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < want) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > PT_SPIN_LIMIT) { atomicOr(err, 1); v = want; break; }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *known = (int)(v - base);
    };
    // steps of the neighbour needed before a chunk's edge values can be read:
    // edge lane (jd=0) at step s uses the neighbour's step s + PT_J - 1
    auto need_for_chunk = [&](int s0, int T_) {
        int n = s0 + PT_C - 1 + T_;
        return n < total ? n : total;
    };

    if (tid == 0) { knownJ = 0; knownK = 0; }
    __syncthreads();
    // first chunk
    if (tid == 0) {
        int kj = knownJ, kk = knownK;
        if (hasJ) wait_for(uJ * ntk + K, need_for_chunk(0, PT_J), &kj);
        if (hasK) wait_for(J * ntk + uK, need_for_chunk(0, PT_K), &kk);
        knownJ = kj; knownK = kk;
    }
    __syncthreads();
    load_own(0, cr, ci, cc);
    load_edge(0, edgeJ, offJ, ejq, eji);
    load_edge(0, edgeK, offK, ekq, eki);

    T prevq = 0, previ = 0;
    for (int s0 = 0; s0 < total; s0 += PT_C) {
        const int s1 = s0 + PT_C;
        // publish progress of the previous chunk and wait for the next one's edges
        if (tid == 0) {
            if (s0 > 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(&prog[myTile], base + (unsigned long long)s0, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            if (s1 < total) {
                int kj = knownJ, kk = knownK;
                if (hasJ) wait_for(uJ * ntk + K, need_for_chunk(s1, PT_J), &kj);
                if (hasK) wait_for(J * ntk + uK, need_for_chunk(s1, PT_K), &kk);
                knownJ = kj; knownK = kk;
            }
        }
        __syncthreads();
        if (s1 < total) {
            load_own(s1, nr, ni_, nc);
            load_edge(s1, edgeJ, offJ, njq, nji);
            load_edge(s1, edgeK, offK, nkq, nki);
        }
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            const int s = s0 + u;
            if (s < total) {
                const int pb = (s + 1) & 1, cb = s & 1;
                if (in_range(s)) {
                    const long c = line + cell_i(s);
                    const uint16_t cd = cc[u];
                    T v = cr[u];
                    const T ic = ci[u];
                    if (FWD) {
                        // column order c-nn, c-ni, c-1
                        int cl = code_dir(cd, dK);
                        if (present_cls(cl)) {
                            T qn = kd > 0 ? lq[pb][kl - 1][jl] : ekq[u];
                            T in_ = kd > 0 ? li[pb][kl - 1][jl] : eki[u];
                            v = v - ((T)code_coef(cl) * in_) * qn;
                        }
                        cl = code_dir(cd, dJ);
                        if (present_cls(cl)) {
                            T qn = jd > 0 ? lq[pb][kl][jl - 1] : ejq[u];
                            T in_ = jd > 0 ? li[pb][kl][jl - 1] : eji[u];
                            v = v - ((T)code_coef(cl) * in_) * qn;
                        }
                        cl = code_dir(cd, dI);
                        if (present_cls(cl)) v = v - ((T)code_coef(cl) * previ) * prevq;
                    } else {
                        // row order c+1, c+ni, c+nn with L_{m,c} = A_{m,c} inv_c
                        int cl = code_dir(cd, dI);
                        if (present_cls(cl)) v = v - ((T)code_coef(cl) * ic) * prevq;
                        cl = code_dir(cd, dJ);
                        if (present_cls(cl)) {
                            T qn = jd > 0 ? lq[pb][kl][jl + 1] : ejq[u];
                            v = v - ((T)code_coef(cl) * ic) * qn;
                        }
                        cl = code_dir(cd, dK);
                        if (present_cls(cl)) {
                            T qn = kd > 0 ? lq[pb][kl + 1][jl] : ekq[u];
                            v = v - ((T)code_coef(cl) * ic) * qn;
                        }
                    }
                    const T q = v * ic;
                    out[c] = q;
                    lq[cb][kl][jl] = q;
                    li[cb][kl][jl] = ic;
                    prevq = q;
                    previ = ic;
                }
                __syncthreads();
            }
        }
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            cr[u] = nr[u]; ci[u] = ni_[u]; cc[u] = nc[u];
            ejq[u] = njq[u]; eji[u] = nji[u]; ekq[u] = nkq[u]; eki[u] = nki[u];
        }
    }
    if (tid == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&prog[myTile], base + (unsigned long long)total, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

static std::atomic<unsigned long long> g_epoch{1};

template <class T>
bool sweep_pipelined(const Geo& g, bool forward, const uint16_t* code, const T* inv, const T* in, T* out,
                     const PcgScalars* S, int* flags, hipStream_t st) {
    const int ntj = (g.nj + PT_J - 1) / PT_J, ntk = (g.nk + PT_K - 1) / PT_K;
    const int tiles = ntj * ntk;
    if (!flags || tiles > 4096) return false;
    // flags layout: [0] error word, [2..] 64-bit progress counters (8-byte aligned)
    unsigned long long* prog = reinterpret_cast<unsigned long long*>(flags + 2);
    unsigned long long ep = g_epoch.fetch_add(1);
    if (forward)
        k_sweep_pipe<T, true><<<tiles, PT_T, 0, st>>>(g, ntj, ntk, code, inv, in, out, S, prog, ep, flags);
    else
        k_sweep_pipe<T, false><<<tiles, PT_T, 0, st>>>(g, ntj, ntk, code, inv, in, out, S, prog, ep, flags);
    return true;
}
int sweep_flag_words(const Geo& g) {
    const int ntj = (g.nj + PT_J - 1) / PT_J, ntk = (g.nk + PT_K - 1) / PT_K;
    return 2 + 2 * ntj * ntk;
}
template bool sweep_pipelined<double>(const Geo&, bool, const uint16_t*, const double*, const double*, double*,
                                      const PcgScalars*, int*, hipStream_t);
template bool sweep_pipelined<float>(const Geo&, bool, const uint16_t*, const float*, const float*, float*,
                                     const PcgScalars*, int*, hipStream_t);
