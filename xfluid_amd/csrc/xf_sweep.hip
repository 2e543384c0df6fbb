// xf_sweep.hip -- pipelined MIC(0) triangular sweeps (persistent line tiles).
//
// The forward sweep q_c = (((r - L q_{c-nn}) - L q_{c-ni}) - L q_{c-1}) inv_c
// (pcg_solver.h:192-204) and the backward sweep (pcg_solver.h:207-220) are
// hyperplane recurrences.  Schedule: the (j,k) plane is cut into 16x16 tiles
// of i-lines; one workgroup owns one tile, one thread one line, and walks i
// with a skew of jl+kl steps (forward; 30-jl-kl backward), so at step s the
// thread's -j / -k neighbours (+j / +k backward) were computed one step
// earlier by the neighbouring threads (exchanged through LDS), its -i
// neighbour by itself.  Tiles form a 2-D pipeline: a tile reads the edge lines
// of its upstream neighbour tiles from global memory once those have published
// enough progress.  Each cell is evaluated with exactly the operations of the
// serial sweep, so the results are bit-identical to the per-hyperplane
// launches and to the reference.
//
// Skewed tile layout.  The sweep inputs live in a per-tile skewed copy:
// element (i, jl, kl) of tile t is stored at ((t*SL + i + jl + kl) * 256 +
// kl*16 + jl), SL = ni + 30 slots per tile, padding zero.  Step s of the
// forward sweep reads slot s of its tile -- one contiguous 256-element row,
// element tid in the lane that uses it -- and step s of the backward sweep
// reads slot ni + 29 - s, again one row.  The forward sweep writes its output
// in the same layout (contiguous rows); the backward sweep writes natural
// order.  Converting r into the layout costs one grid pass per application.
//
// Per chunk of C steps the step loop touches only LDS and registers.  At the
// chunk top the workgroup waits for its own stores of the chunk before last,
// publishes that progress, waits for the neighbour progress the next chunk's
// edge values need, issues the next chunk's loads (registers, consumed a chunk
// later) and the previous chunk's stores.  Edge values that other tiles read
// are stored and loaded with device-scope atomic accesses.  Progress counters
// carry an epoch in their high bits (no reset between sweeps).  Waits are
// bounded: on timeout the kernel records an error and finishes, and the host
// redoes the solve with per-hyperplane launches.
#include "xf_internal.h"

#define PT_J 16
#define PT_K 16
#define PT_T (PT_J * PT_K)
#define PT_PAD (PT_J + PT_K - 2)
#define PT_SPIN_LIMIT (1 << 20)  // polls (~0.5 s) before a wait is declared stuck

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

// LDS-only barrier: orders the LDS exchange between steps without waiting
// for outstanding global loads/stores (a full __syncthreads would).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// natural -> skewed copy; padding slots and lanes outside the grid get zero
template <class D, class Sx>
__global__ void k_to_skewed(Geo g, int ntk, long total_sk, const Sx* __restrict__ src, D* __restrict__ dst,
                            const PcgScalars* S) {
    if (S && S->done) return;
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= total_sk) return;
    const int SL = g.ni + PT_PAD;
    const long tile = e / ((long)SL * PT_T);
    const int rem = (int)(e - tile * (long)SL * PT_T);
    const int slot = rem / PT_T, lane = rem % PT_T;
    const int jl = lane % PT_J, kl = lane / PT_J;
    const int j = (int)(tile / ntk) * PT_J + jl, k = (int)(tile % ntk) * PT_K + kl;
    const int i = slot - jl - kl;
    D v = (D)0;
    if (j < g.nj && k < g.nk && i >= 0 && i < g.ni) v = (D)src[cidx(g, i, j, k)];
    dst[e] = v;
}

template <class T, bool FWD, int PT_C>
__global__ __launch_bounds__(PT_T) void k_sweep_pipe(Geo g, int ntj, int ntk, const uint16_t* __restrict__ code_sk,
                                                     const T* __restrict__ inv_sk, const T* __restrict__ in_sk,
                                                     T* __restrict__ out, const PcgScalars* S,
                                                     unsigned long long* prog, unsigned long long epoch, int* err) {
    __shared__ T lq[2 * PT_C][PT_K][PT_J];  // q ring: exchange between steps + output staging
    __shared__ T li[2][PT_K][PT_J];         // pivots of the -j/-k neighbours (forward sweep)
    __shared__ int knownJ, knownK;
    if (S && S->done) return;
    int J, K;
    tile_of_block(blockIdx.x, ntj, ntk, FWD, &J, &K);
    const int tid = threadIdx.x, jl = tid % PT_J, kl = tid / PT_J;
    const int j = J * PT_J + jl, k = K * PT_K + kl;
    const bool valid = j < g.nj && k < g.nk;
    const int jd = FWD ? jl : (PT_J - 1 - jl), kd = FWD ? kl : (PT_K - 1 - kl);
    const int skew = jd + kd;
    const int total = g.ni + PT_PAD;  // steps per sweep = slots per tile
    const int nchunks = (total + PT_C - 1) / PT_C;
    const int uJ = FWD ? J - 1 : J + 1, uK = FWD ? K - 1 : K + 1;
    const bool hasJ = uJ >= 0 && uJ < ntj, hasK = uK >= 0 && uK < ntk;
    const int myTile = J * ntk + K;
    const unsigned long long base = epoch << 20;
    const bool edgeJ = (jd == 0) && hasJ, edgeK = (kd == 0) && hasK;  // reads a neighbour tile
    const bool sharedOut = (jd == PT_J - 1) || (kd == PT_K - 1);       // read by a neighbour tile
    const long tileRow0 = (long)myTile * total * PT_T;
    const int dJ = FWD ? D_YM : D_YP, dK = FWD ? D_ZM : D_ZP, dI = FWD ? D_XM : D_XP;
    const long line = valid ? (long)g.ni * j + g.nn * k : 0;
    // forward edge sources: cell (i, j-1, k) is lane (15, kl) of tile J-1 and
    // cell (i, j, k-1) lane (jl, 15) of tile K-1, both at slot s + 15
    const long upJ = hasJ ? (long)(uJ * ntk + K) * total * PT_T + kl * PT_J + (PT_J - 1) : 0;
    const long upK = hasK ? (long)(J * ntk + uK) * total * PT_T + (PT_K - 1) * PT_J + jl : 0;

    auto slot_of = [&](int s) { return FWD ? s : total - 1 - s; };
    auto in_range = [&](int s) { const int il = s - skew; return valid && il >= 0 && il < g.ni; };
    auto nat_of = [&](int s) { const int il = s - skew; return line + (FWD ? il : g.ni - 1 - il); };

    T cr[PT_C], ci[PT_C], ejq[PT_C], eji[PT_C], ekq[PT_C], eki[PT_C];
    T nr[PT_C], nI[PT_C], njq[PT_C], nji[PT_C], nkq[PT_C], nki[PT_C];
    uint32_t cc[PT_C], nc[PT_C];

    auto load_chunk = [&](int s0, T* r_, T* i_, uint32_t* c_, T* jq, T* ji, T* kq, T* ki) {
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            const int s = s0 + u;
            const bool live = s < total;
            const long e = tileRow0 + (long)slot_of(live ? s : 0) * PT_T + tid;
            r_[u] = live ? in_sk[e] : (T)0;
            i_[u] = live ? inv_sk[e] : (T)0;
            c_[u] = live ? (uint32_t)code_sk[e] : 0u;
            const bool ok = live && in_range(s);
            jq[u] = (T)0; ji[u] = (T)0; kq[u] = (T)0; ki[u] = (T)0;
            if (FWD) {
                const long sj = upJ + (long)(s + PT_J - 1) * PT_T, sk = upK + (long)(s + PT_K - 1) * PT_T;
                if (ok && edgeJ) {
                    jq[u] = __hip_atomic_load(&out[sj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ji[u] = inv_sk[sj];
                }
                if (ok && edgeK) {
                    kq[u] = __hip_atomic_load(&out[sk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ki[u] = inv_sk[sk];
                }
            } else {
                const long c = nat_of(s);
                if (ok && edgeJ) jq[u] = __hip_atomic_load(&out[c + g.ni], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (ok && edgeK) kq[u] = __hip_atomic_load(&out[c + g.nn], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    };
    // outputs of a chunk: forward -> skewed rows (coalesced), backward -> natural cells
    auto flush_chunk = [&](int s0) {
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            const int s = s0 + u;
            if (s < 0 || s >= total || !in_range(s)) continue;
            const T v = lq[s % (2 * PT_C)][kl][jl];
            const long e = FWD ? tileRow0 + (long)s * PT_T + tid : nat_of(s);
            if (sharedOut) __hip_atomic_store(&out[e], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else out[e] = v;
        }
    };
    auto wait_for = [&](int t, int need, int* known) {
        if (*known >= need) return;
        const unsigned long long want = base + (unsigned long long)need;
        unsigned long long v;
        int spins = 0;
        while ((v = __hip_atomic_load(&prog[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < want) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > PT_SPIN_LIMIT) { atomicOr(err, 1); v = want; break; }
        }
        *known = (int)(v - base);
    };
    // an edge lane at step s uses the neighbour's step s + 15; progress counts completed steps
    auto need_for_chunk = [&](int s0, int TT) {
        const int n = s0 + PT_C - 1 + TT;
        return n < total ? n : total;
    };
    auto wait_chunk = [&](int s0) {
        if (tid == 0) {
            int kj = knownJ, kk = knownK;
            if (hasJ) wait_for(uJ * ntk + K, need_for_chunk(s0, PT_J), &kj);
            if (hasK) wait_for(J * ntk + uK, need_for_chunk(s0, PT_K), &kk);
            knownJ = kj; knownK = kk;
        }
    };

    if (tid == 0) { knownJ = 0; knownK = 0; }
    __syncthreads();
    wait_chunk(0);
    __syncthreads();
    load_chunk(0, cr, ci, cc, ejq, eji, ekq, eki);

    T prevq = 0, previ = 0;
    for (int t = 0; t < nchunks; ++t) {
        const int s0 = t * PT_C, s1 = s0 + PT_C;
        // everything issued before this point is complete -> the stores of
        // chunk t-2 (issued at the top of chunk t-1) are visible: publish them
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __syncthreads();
        if (tid == 0 && t >= 2)
            __hip_atomic_store(&prog[myTile], base + (unsigned long long)(s0 - PT_C), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (s1 < total) wait_chunk(s1);
        __syncthreads();
        if (s1 < total) load_chunk(s1, nr, nI, nc, njq, nji, nkq, nki);
        if (t >= 1) flush_chunk(s0 - PT_C);
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            const int s = s0 + u;
            if (s < total) {
                const int pb = (s + 1) & 1, cb = s & 1;
                const int rp = (s + 2 * PT_C - 1) % (2 * PT_C), rc = s % (2 * PT_C);
                if (in_range(s)) {
                    const uint16_t cd = (uint16_t)cc[u];
                    T v = cr[u];
                    const T ic = ci[u];
                    // each update is computed unconditionally and kept only when the
                    // entry is present, which leaves the serial arithmetic unchanged
                    const int ck = code_dir(cd, dK), cj = code_dir(cd, dJ), cx = code_dir(cd, dI);
                    const T ak = (T)code_coef(ck), aj = (T)code_coef(cj), ai = (T)code_coef(cx);
                    if (FWD) {
                        // column order c-nn, c-ni, c-1; L_{c,m} = A_{c,m} inv_m
                        const T qk = kd > 0 ? lq[rp][kl - 1][jl] : ekq[u];
                        const T ik = kd > 0 ? li[pb][kl - 1][jl] : eki[u];
                        const T qj = jd > 0 ? lq[rp][kl][jl - 1] : ejq[u];
                        const T ij = jd > 0 ? li[pb][kl][jl - 1] : eji[u];
                        T w = v - (ak * ik) * qk;
                        v = present_cls(ck) ? w : v;
                        w = v - (aj * ij) * qj;
                        v = present_cls(cj) ? w : v;
                        w = v - (ai * previ) * prevq;
                        v = present_cls(cx) ? w : v;
                    } else {
                        // row order c+1, c+ni, c+nn; L_{m,c} = A_{m,c} inv_c
                        const T qj = jd > 0 ? lq[rp][kl][jl + 1] : ejq[u];
                        const T qk = kd > 0 ? lq[rp][kl + 1][jl] : ekq[u];
                        T w = v - (ai * ic) * prevq;
                        v = present_cls(cx) ? w : v;
                        w = v - (aj * ic) * qj;
                        v = present_cls(cj) ? w : v;
                        w = v - (ak * ic) * qk;
                        v = present_cls(ck) ? w : v;
                    }
                    const T q = v * ic;
                    lq[rc][kl][jl] = q;
                    if (FWD) li[cb][kl][jl] = ic;
                    prevq = q;
                    previ = ic;
                }
                lds_barrier();
            }
        }
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            cr[u] = nr[u]; ci[u] = nI[u]; cc[u] = nc[u];
            ejq[u] = njq[u]; eji[u] = nji[u]; ekq[u] = nkq[u]; eki[u] = nki[u];
        }
    }
    // flush the last chunk and publish completion
    __syncthreads();
    flush_chunk((nchunks - 1) * PT_C);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (tid == 0)
        __hip_atomic_store(&prog[myTile], base + (unsigned long long)total, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// I/O-wave schedule.  The workgroup has the 4 compute waves of the tile plus a
// load wave and a store wave that perform every global memory access.  Compute waves work on
// LDS only: at step s they read their inputs from an LDS ring slot s % R and
// the neighbours' results of step s-1 from the output ring, write q to the
// output ring, and meet at one LDS barrier per step.  Each memory wave runs one
// software-pipelined iteration per step between the barriers.  The load wave
// moves the inputs of step s+2 (loaded L iterations earlier) into the ring,
// polls the upstream tiles' counters once per L iterations and issues the edge
// and own-input loads of step s+2+L.  The store wave stores the outputs of step
// s-1 and publishes the progress count of steps whose device-scope edge stores
// are known complete (explicit vmcnt wait leaving the last P-1 iterations in
// flight).  Two memory waves double the number of memory instructions that can
// be in flight (a wave's vmcnt counter holds at most 63).  Only when an upstream
// tile has not yet published what the loads need does the load wave wait
// (bounded), which holds the compute waves at the barrier.
// Both sweeps read and write the skewed layout; the backward result is
// converted to natural order afterwards.
#define IO_R 16  // LDS ring depth in steps

template <class T>
struct IoRegs {
    T r[4], i[4];
    uint64_t c;  // four 16-bit stencil codes
    T e;         // one edge value (lanes 0-15 j-edge q, 16-31 k-edge q, 32-63 their pivots)
};

template <class T, bool FWD, int L, int P>
__global__ __launch_bounds__(PT_T + 128) void k_sweep_io(Geo g, int ntj, int ntk, const uint16_t* __restrict__ code_sk,
                                                         const T* __restrict__ inv_sk, const T* __restrict__ in_sk,
                                                         T* __restrict__ out_sk, T* __restrict__ eb,
                                                         const PcgScalars* S, unsigned long long* prog,
                                                         unsigned long long epoch, int* err) {
    __shared__ T s_r[IO_R][PT_T], s_i[IO_R][PT_T], s_q[IO_R][PT_T];
    __shared__ uint16_t s_c[IO_R][PT_T];
    __shared__ T s_e[IO_R][64];
    if (S && S->done) return;
    int J, K;
    tile_of_block(blockIdx.x, ntj, ntk, FWD, &J, &K);
    const int tid = threadIdx.x;
    const int total = g.ni + PT_PAD;
    const int myTile = J * ntk + K;
    const long tileRow0 = (long)myTile * total * PT_T;
    const unsigned long long base = epoch << 20;
    const int uJ = FWD ? J - 1 : J + 1, uK = FWD ? K - 1 : K + 1;
    const bool hasJ = uJ >= 0 && uJ < ntj, hasK = uK >= 0 && uK < ntk;
    auto slot_of = [&](int s) { return FWD ? s : total - 1 - s; };

    if (tid < PT_T) {
        // ------------------------------ compute waves ------------------------------
        // The inputs of step s (own value, pivot, code, edge values, and for the
        // forward sweep the neighbours' pivots) are read one step ahead; only the
        // neighbours' results of step s-1 are read after the barrier.
        const int jl = tid % PT_J, kl = tid / PT_J;
        const int j = J * PT_J + jl, k = K * PT_K + kl;
        const bool valid = j < g.nj && k < g.nk;
        const int jd = FWD ? jl : (PT_J - 1 - jl), kd = FWD ? kl : (PT_K - 1 - kl);
        const int skew = jd + kd;
        const int dJ = FWD ? D_YM : D_YP, dK = FWD ? D_ZM : D_ZP, dI = FWD ? D_XM : D_XP;
        // lanes of the j / k neighbours (own lane at the tile edge: the value comes from the edge ring)
        const int lj = jd > 0 ? tid + (FWD ? -1 : 1) : tid, lk = kd > 0 ? tid + (FWD ? -PT_J : PT_J) : tid;
        // Everything but the neighbours' results is known a step ahead: the own
        // value and pivot, the edge values, and the three products A_cm * inv
        // (forward: the neighbour's pivot, backward: the own pivot) with their
        // presence flags.  Each product is formed exactly as in the serial sweep.
        struct In {
            T r, ic, ej, ek, ck, cj, ci;
            int pk, pj, pi;
        };
        auto fetch = [&](int s, T prev_ic) {
            In x;
            const int sl = s & (IO_R - 1), pv = (s + IO_R - 1) & (IO_R - 1);
            x.r = s_r[sl][tid];
            x.ic = s_i[sl][tid];
            const uint16_t cd = s_c[sl][tid];
            x.ej = s_e[sl][kl];
            x.ek = s_e[sl][16 + jl];
            const int ck = code_dir(cd, dK), cj = code_dir(cd, dJ), cx = code_dir(cd, dI);
            const T ak = (T)code_coef(ck), aj = (T)code_coef(cj), ai = (T)code_coef(cx);
            x.pk = present_cls(ck);
            x.pj = present_cls(cj);
            x.pi = present_cls(cx);
            if (FWD) {
                const T ij = jd > 0 ? s_i[pv][lj] : s_e[sl][32 + kl];
                const T ik = kd > 0 ? s_i[pv][lk] : s_e[sl][48 + jl];
                x.ck = ak * ik;
                x.cj = aj * ij;
                x.ci = ai * prev_ic;
            } else {
                x.ck = ak * x.ic;
                x.cj = aj * x.ic;
                x.ci = ai * x.ic;
            }
            return x;
        };
        __syncthreads();  // prologue: steps 0 and 1 are in the ring
        In cur = fetch(0, (T)0);
        T prevq = 0;
        for (int s = 0; s < total; ++s) {
            const int sl = s & (IO_R - 1), pv = (s + IO_R - 1) & (IO_R - 1);
            // results of step s-1 (the barrier of the last step made them visible)
            const T qjl = s_q[pv][lj], qkl = s_q[pv][lk];
            // issue these two reads first: they are on the critical path, the prefetch is not
            __asm__ volatile("" ::: "memory");
            // slot s+1 was filled two iterations ago; the -i neighbour of step s+1 is this lane's step s
            const In nxt = fetch(s + 1 < total ? s + 1 : s, cur.ic);
            const int il = s - skew;
            if (valid && il >= 0 && il < g.ni) {
                const T qk = kd > 0 ? qkl : cur.ek;
                const T qj = jd > 0 ? qjl : cur.ej;
                T v = cur.r, w;
                if (FWD) {
                    // column order c-nn, c-ni, c-1
                    w = v - cur.ck * qk;
                    v = cur.pk ? w : v;
                    w = v - cur.cj * qj;
                    v = cur.pj ? w : v;
                    w = v - cur.ci * prevq;
                    v = cur.pi ? w : v;
                } else {
                    // row order c+1, c+ni, c+nn
                    w = v - cur.ci * prevq;
                    v = cur.pi ? w : v;
                    w = v - cur.cj * qj;
                    v = cur.pj ? w : v;
                    w = v - cur.ck * qk;
                    v = cur.pk ? w : v;
                }
                const T q = v * cur.ic;
                s_q[sl][tid] = q;
                prevq = q;
            }
            cur = nxt;
            lds_barrier();
        }
        return;
    }

    const int lane = tid & 63;
    const long guard = (long)ntj * ntk * total * PT_T;  // one spare row after the tiles
    if (tid >= PT_T + 64) {
        // ------------------------------- store wave -------------------------------
        // Outputs of step s-1 at iteration s: plain coalesced stores of the whole
        // row (read by the next kernel; lane l writes elements 2l, 2l+1 and
        // 128+2l, 129+2l), and the 31 tile-edge values once more, with
        // device-scope stores, into the tile's edge buffer (64 entries per slot:
        // 0-15 the edge column by kl, 16-31 the edge row by jl), which the
        // neighbour tiles read during this kernel.  Progress is published P
        // iterations later, when an explicit vmcnt wait has shown those stores
        // complete.
        const int ecol = FWD ? (PT_J - 1) : 0;  // edge column / row of this tile's outputs
        const int eel = lane < 16 ? lane * PT_J + ecol : lane < 32 ? ecol * PT_J + (lane - 16) : 0;
        const long ebTile = (long)myTile * total * 64;
        const long ebGuard = (long)ntj * ntk * total * 64;
        constexpr int QS = 2 + 1 + 1;  // 2 row stores, 1 edge store, 1 publish
        constexpr int NW = (P - 1) * QS;
        constexpr int WAITCNT = (NW & 15) | (7 << 4) | (15 << 8) | ((NW >> 4) << 14);
        static_assert(NW < 64, "vmcnt field");
        typedef HIP_vector_type<T, 2> T2;
        auto store_out = [&](int y) {
            const bool ok = y >= 0 && y < total;
            const int sl = (y + IO_R) & (IO_R - 1);
            const long rowb = ok ? tileRow0 + (long)slot_of(y) * PT_T : guard;
            const T2 a = *(const T2*)&s_q[sl][2 * lane];
            const T2 b = *(const T2*)&s_q[sl][PT_T / 2 + 2 * lane];
            const T ev = s_q[sl][eel];
            *(T2*)(out_sk + rowb + 2 * lane) = a;
            *(T2*)(out_sk + rowb + PT_T / 2 + 2 * lane) = b;
            const long edst = ok ? ebTile + (long)slot_of(y) * 64 + lane : ebGuard + lane;
            __hip_atomic_store(eb + edst, ev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        };
        __syncthreads();
        for (int s = 0; s < total; ++s) {
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_waitcnt(WAITCNT);  // stores of iteration s-P are complete
            __builtin_amdgcn_sched_barrier(0);
            {
                const int done = s < P ? 0 : s - P;  // steps < s-P
                __hip_atomic_store(prog + myTile, base + (unsigned long long)done, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            __builtin_amdgcn_sched_barrier(0);
            store_out(s - 1);
            __builtin_amdgcn_sched_barrier(0);
            lds_barrier();
        }
        store_out(total - 1);
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_store(prog + myTile, base + (unsigned long long)total, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        return;
    }

    // -------------------------------- load wave --------------------------------
    const long upJrow = hasJ ? (long)(uJ * ntk + K) * total * PT_T : tileRow0;
    const long upKrow = hasK ? (long)(J * ntk + uK) * total * PT_T : tileRow0;
    // edge value of this lane: lanes 0-31 the upstream tiles' results from their
    // edge buffers, lanes 32-63 the matching pivots from the skewed pivot array
    const int el = lane & 31;
    const bool eJ = el < 16;
    const int ucol = FWD ? (PT_J - 1) : 0;  // the upstream tile's edge column / row
    const bool eHas = eJ ? hasJ : hasK;
    const long ebJ = hasJ ? (long)(uJ * ntk + K) * total * 64 : 0;
    const long ebK = hasK ? (long)(J * ntk + uK) * total * 64 : 0;
    const long eoff = lane < 32 ? (eJ ? ebJ + el : ebK + el)
                                : (eJ ? (upJrow + (long)el * PT_J + ucol) : (upKrow + (long)ucol * PT_J + (el - 16)));
    const long estride = lane < 32 ? 64 : PT_T;
    const T* esrc = lane < 32 ? (const T*)eb : inv_sk;
    unsigned long long* pJ = prog + (hasJ ? uJ * ntk + K : myTile);
    unsigned long long* pK = prog + (hasK ? J * ntk + uK : myTile);
    const unsigned long long needHas = (hasJ ? 1ull : 0ull) | (hasK ? 2ull : 0ull);
    int known = 0;  // upstream steps known complete (min over the upstream tiles)

    auto refresh = [&](unsigned long long vj, unsigned long long vk) {
        int a = hasJ ? (int)(vj >= base ? vj - base : 0) : total;
        int b = hasK ? (int)(vk >= base ? vk - base : 0) : total;
        int m = a < b ? a : b;
        known = m > known ? m : known;
    };
    auto wait_upstream = [&](int need) {
        if (!needHas || known >= need) return;
        int spins = 0;
        while (known < need) {
            const unsigned long long vj = __hip_atomic_load(pJ, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long vk = __hip_atomic_load(pK, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            refresh(vj, vk);
            if (known >= need) break;
            __builtin_amdgcn_s_sleep(1);
            if (++spins > PT_SPIN_LIMIT) {
                if (lane == 0) atomicOr(err, 1);
                known = total;
            }
        }
    };
    // an edge lane at step y reads the upstream tile's step y + 15 (count y + 16)
    auto edge_need = [&](int y) { int n = y + PT_J; return n < total ? n : total; };
    // lane l moves elements 2l, 2l+1 and 128+2l, 129+2l of each row (coalesced)
    typedef HIP_vector_type<T, 2> T2;
    auto issue_loads = [&](int y, IoRegs<T>& R) {
        const int yy = y < total ? y : total - 1;
        const long row = tileRow0 + (long)slot_of(yy) * PT_T + 2 * lane;
        const T2 r0 = *(const T2*)(in_sk + row), r1 = *(const T2*)(in_sk + row + PT_T / 2);
        const T2 i0 = *(const T2*)(inv_sk + row), i1 = *(const T2*)(inv_sk + row + PT_T / 2);
        R.r[0] = r0.x; R.r[1] = r0.y; R.r[2] = r1.x; R.r[3] = r1.y;
        R.i[0] = i0.x; R.i[1] = i0.y; R.i[2] = i1.x; R.i[3] = i1.y;
        const uint32_t c0 = *(const uint32_t*)(code_sk + row), c1 = *(const uint32_t*)(code_sk + row + PT_T / 2);
        R.c = (uint64_t)c0 | ((uint64_t)c1 << 32);
        const int es = FWD ? yy + PT_J - 1 : total - PT_J - yy;  // upstream slot of the edge cells
        // without an upstream value the lane reads a harmless in-range element of its source
        const long ea = (eHas && es >= 0 && es < total) ? eoff + (long)es * estride : (lane < 32 ? (long)lane : row);
        R.e = __hip_atomic_load(esrc + ea, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto put_ring = [&](int y, const IoRegs<T>& R) {
        const int sl = y & (IO_R - 1);
        *(T2*)&s_r[sl][2 * lane] = T2{R.r[0], R.r[1]};
        *(T2*)&s_r[sl][PT_T / 2 + 2 * lane] = T2{R.r[2], R.r[3]};
        *(T2*)&s_i[sl][2 * lane] = T2{R.i[0], R.i[1]};
        *(T2*)&s_i[sl][PT_T / 2 + 2 * lane] = T2{R.i[2], R.i[3]};
        *(uint32_t*)&s_c[sl][2 * lane] = (uint32_t)R.c;
        *(uint32_t*)&s_c[sl][PT_T / 2 + 2 * lane] = (uint32_t)(R.c >> 32);
        s_e[sl][lane] = R.e;
    };

    // prologue: steps 0 and 1 into the ring, steps 2..L+1 in flight
    wait_upstream(edge_need(L + 1));
    {
        IoRegs<T> R0, R1;
        issue_loads(0, R0);
        issue_loads(1, R1);
        put_ring(0, R0);
        put_ring(1, R1);
    }
    IoRegs<T> ring[L];
#pragma unroll
    for (int u = 0; u < L; ++u) issue_loads(u + 2, ring[(u + 2) % L]);
    __syncthreads();

    for (int s0 = 0; s0 < total; s0 += L) {
        // edges of the steps loaded in this trip (s0+2+L .. s0+2L+1)
        if (s0 + 2 + L < total) wait_upstream(edge_need(s0 + 2 * L + 1));
        unsigned long long pjv = 0, pkv = 0;  // polled at the first iteration, read at the last
#pragma unroll
        for (int u = 0; u < L; ++u) {
            const int s = s0 + u;
            // inputs of step s+2 (ring slots beyond the last step are never read)
            put_ring(s + 2, ring[(u + 2) % L]);
            // upstream progress: polled at the first iteration of the trip, read at the last
            if (u == 0) {
                pjv = __hip_atomic_load(pJ, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                pkv = __hip_atomic_load(pK, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (u == L - 1) refresh(pjv, pkv);
            // loads of step s+2+L
            issue_loads(s + 2 + L, ring[(u + 2) % L]);
            if (s < total) lds_barrier();
        }
    }
}

// skewed -> natural (the backward sweep's output)
template <class T>
__global__ void k_from_skewed(Geo g, int ntk, const T* __restrict__ src, T* __restrict__ dst, const PcgScalars* S) {
    if (S && S->done) return;
    const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const int i = (int)(c % g.ni), j = (int)((c / g.ni) % g.nj), k = (int)(c / g.nn);
    const int SL = g.ni + PT_PAD;
    const long tile = (long)(j / PT_J) * ntk + k / PT_K;
    const int jl = j % PT_J, kl = k % PT_K;
    dst[c] = src[(tile * SL + i + jl + kl) * PT_T + kl * PT_J + jl];
}

static void tiles_of(const Geo& g, int* ntj, int* ntk) {
    *ntj = (g.nj + PT_J - 1) / PT_J;
    *ntk = (g.nk + PT_K - 1) / PT_K;
}
long sweep_skewed_size(const Geo& g) {
    int ntj, ntk;
    tiles_of(g, &ntj, &ntk);
    return (long)ntj * ntk * (g.ni + PT_PAD) * PT_T + PT_T;  // + one guard row
}
// Flags layout (ints): [0] error word, [1] pad, [2, 2+2T) 64-bit progress counters.
int sweep_flag_words(const Geo& g) {
    int ntj, ntk;
    tiles_of(g, &ntj, &ntk);
    return 2 + 2 * ntj * ntk;
}

template <class T>
void sweep_prepare(const Geo& g, const uint16_t* code, const T* inv, const SweepSkew<T>& sk, hipStream_t st) {
    int ntj, ntk;
    tiles_of(g, &ntj, &ntk);
    const long n = sweep_skewed_size(g);
    if (code) k_to_skewed<uint16_t, uint16_t><<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, ntk, n, code, sk.code, nullptr);
    k_to_skewed<T, T><<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, ntk, n, inv, sk.inv, nullptr);
}

template <class T, int C>
static void launch_pair(const Geo& g, int ntj, int ntk, const SweepSkew<T>& sk, T* z, const PcgScalars* S,
                        unsigned long long* prog, unsigned long long e1, unsigned long long e2, int* flags,
                        hipStream_t st) {
    const int tiles = ntj * ntk;
    k_sweep_pipe<T, true, C><<<tiles, PT_T, 0, st>>>(g, ntj, ntk, sk.code, sk.inv, sk.r, sk.q, S, prog, e1, flags);
    k_sweep_pipe<T, false, C><<<tiles, PT_T, 0, st>>>(g, ntj, ntk, sk.code, sk.inv, sk.q, z, S, prog, e2, flags);
}

// z = (L L^T)^{-1} r with the pipelined schedule; sweep_kernel 2, 3, 4, 5 ->
// chunks of 8, 4, 2, 1 steps; 6, 7 -> memory-wave schedule with load lead 6, 8
// and publication lag 8, 10 steps.  Returns false when the grid does not fit the
// progress-counter encoding (the caller then uses per-hyperplane launches).
// t0/t1 (optional) are recorded around the two sweep launches.
template <class T>
bool sweep_precond_pipelined(const Geo& g, int kind, const SweepSkew<T>& sk, const T* r, T* z, const PcgScalars* S,
                             int* flags, unsigned long long* epoch, hipEvent_t t0, hipEvent_t t1, hipStream_t st) {
    int ntj, ntk;
    tiles_of(g, &ntj, &ntk);
    const int tiles = ntj * ntk;
    if (!flags || !epoch || tiles > 4096 || g.ni + PT_PAD >= (1 << 20)) return false;
    unsigned long long* prog = reinterpret_cast<unsigned long long*>(flags + 2);
    const long n = sweep_skewed_size(g);
    k_to_skewed<T, T><<<xf_blocks(n), XF_BLOCK, 0, st>>>(g, ntk, n, r, sk.r, S);
    const unsigned long long e1 = (*epoch)++, e2 = (*epoch)++;
    if (t0) hipEventRecord(t0, st);
    if (kind == 6 || kind == 7) {
        if (!sk.z || !sk.eb) return false;
        if (kind == 6) {
            k_sweep_io<T, true, 6, 8><<<tiles, PT_T + 128, 0, st>>>(g, ntj, ntk, sk.code, sk.inv, sk.r, sk.q, sk.eb, S, prog, e1, flags);
            k_sweep_io<T, false, 6, 8><<<tiles, PT_T + 128, 0, st>>>(g, ntj, ntk, sk.code, sk.inv, sk.q, sk.z, sk.eb, S, prog, e2, flags);
        } else {
            k_sweep_io<T, true, 8, 10><<<tiles, PT_T + 128, 0, st>>>(g, ntj, ntk, sk.code, sk.inv, sk.r, sk.q, sk.eb, S, prog, e1, flags);
            k_sweep_io<T, false, 8, 10><<<tiles, PT_T + 128, 0, st>>>(g, ntj, ntk, sk.code, sk.inv, sk.q, sk.z, sk.eb, S, prog, e2, flags);
        }
        if (t1) hipEventRecord(t1, st);
        k_from_skewed<T><<<xf_blocks(g.N), XF_BLOCK, 0, st>>>(g, ntk, sk.z, z, S);
        return true;
    }
    switch (kind) {
        case 2: launch_pair<T, 8>(g, ntj, ntk, sk, z, S, prog, e1, e2, flags, st); break;
        case 4: launch_pair<T, 2>(g, ntj, ntk, sk, z, S, prog, e1, e2, flags, st); break;
        case 5: launch_pair<T, 1>(g, ntj, ntk, sk, z, S, prog, e1, e2, flags, st); break;
        default: launch_pair<T, 4>(g, ntj, ntk, sk, z, S, prog, e1, e2, flags, st); break;
    }
    if (t1) hipEventRecord(t1, st);
    return true;
}
template void sweep_prepare<double>(const Geo&, const uint16_t*, const double*, const SweepSkew<double>&, hipStream_t);
template void sweep_prepare<float>(const Geo&, const uint16_t*, const float*, const SweepSkew<float>&, hipStream_t);
template bool sweep_precond_pipelined<double>(const Geo&, int, const SweepSkew<double>&, const double*, double*,
                                              const PcgScalars*, int*, unsigned long long*, hipEvent_t, hipEvent_t,
                                              hipStream_t);
template bool sweep_precond_pipelined<float>(const Geo&, int, const SweepSkew<float>&, const float*, float*,
                                             const PcgScalars*, int*, unsigned long long*, hipEvent_t, hipEvent_t,
                                             hipStream_t);
