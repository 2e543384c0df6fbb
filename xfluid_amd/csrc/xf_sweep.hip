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
#include "xf_internal.h"

// Tile PT_J x PT_K lines, chunks of PT_C steps (loads one chunk ahead,
// stores one chunk behind); template parameters of the kernel.
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

// Per chunk of PT_C steps the step loop touches only LDS and registers.  At
// the chunk top the workgroup (1) waits for its own earlier global stores,
// (2) publishes progress, (3) waits for the neighbour tiles' progress that the
// next chunk's edge values need, (4) issues the next chunk's loads and (5)
// flushes the previous chunk's outputs.  Loads and stores are cooperative:
// consecutive lanes cover consecutive i of one line, so a wave touches a few
// 64-byte segments per instruction instead of one cache line per lane; the
// loaded values reach LDS at the end of the chunk, when they have arrived.
// Edge lines that other tiles read are stored and loaded with device-scope
// (sc1) atomic accesses, so the hand-off needs no cache-wide fences.
// Transport of the tile hand-off (edge lines + progress counters):
//  FENCE = false: edge values and counters with device-scope (sc1) atomic
//                 loads/stores, no fences;
//  FENCE = true : plain loads/stores of the edge values; the producer releases
//                 at agent scope before publishing a counter, the consumer
//                 acquires at agent scope after reading one (HIP memory model).
template <class T, bool FENCE>
__device__ inline T load_shared(const T* p) {
    if constexpr (FENCE) return *p;
    else return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T, bool FENCE>
__device__ inline void store_shared(T* p, T v) {
    if constexpr (FENCE) *p = v;
    else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T, bool FWD, int PT_J, int PT_K, int PT_C, bool FENCE>
__device__ __forceinline__ void sweep_tile(const int J, const int K, const Geo& g, int ntj, int ntk,
                                           const uint16_t* __restrict__ code, const T* __restrict__ inv,
                                           const T* __restrict__ in, T* __restrict__ out,
                                           unsigned long long* prog, unsigned long long epoch, int* err) {
    constexpr int NT = PT_J * PT_K;
    constexpr int EL = NT * PT_C / NT;      // elements per lane in a cooperative chunk load (= PT_C)
    __shared__ T sr[2][PT_C][NT];           // own input (r or q) of the chunk, per line
    __shared__ T si[2][PT_C][NT];           // own pivots
    __shared__ uint32_t sc[2][PT_C][NT];    // own stencil codes
    __shared__ T ejq[2][PT_C][PT_K], eji[2][PT_C][PT_K];  // -j (+j) neighbour tile edge line
    __shared__ T ekq[2][PT_C][PT_J], eki[2][PT_C][PT_J];  // -k (+k) neighbour tile edge line
    __shared__ T lq[2 * PT_C][PT_K][PT_J];  // q ring: exchange between steps + output staging
    __shared__ T li[2][PT_K][PT_J];         // pivots of the -j/-k neighbours (forward sweep)
    __shared__ int knownJ, knownK;
    const int tid = threadIdx.x, jl = tid % PT_J, kl = tid / PT_J;
    const int jd = FWD ? jl : (PT_J - 1 - jl), kd = FWD ? kl : (PT_K - 1 - kl);
    const int total = g.ni + (PT_J - 1) + (PT_K - 1);
    const int nchunks = (total + PT_C - 1) / PT_C;
    const int uJ = FWD ? J - 1 : J + 1, uK = FWD ? K - 1 : K + 1;
    const bool hasJ = uJ >= 0 && uJ < ntj, hasK = uK >= 0 && uK < ntk;
    const int myTile = J * ntk + K;
    const unsigned long long base = epoch << 20;
    const long offJ = FWD ? -(long)g.ni : (long)g.ni;
    const long offK = FWD ? -g.nn : g.nn;
    const int dJ = FWD ? D_YM : D_YP, dK = FWD ? D_ZM : D_ZP, dI = FWD ? D_XM : D_XP;
    const int skew = jd + kd;
    const bool valid = (J * PT_J + jl) < g.nj && (K * PT_K + kl) < g.nk;

    // line L of the tile: (jl, kl) = (L % PT_J, L / PT_J)
    auto line_info = [&](int L, long* base_c, int* sk, bool* ok) {
        const int a = L % PT_J, b = L / PT_J;
        const int jj = J * PT_J + a, kk = K * PT_K + b;
        *ok = jj < g.nj && kk < g.nk;
        *base_c = *ok ? (long)g.ni * jj + g.nn * kk : 0;
        *sk = FWD ? (a + b) : ((PT_J - 1 - a) + (PT_K - 1 - b));
    };
    auto cell_of = [&](int s, int sk, int* i) {
        const int il = s - sk;
        *i = FWD ? il : g.ni - 1 - il;
        return il >= 0 && il < g.ni;
    };

    // cooperative chunk load into registers (element e = tid + NT*m: line e / PT_C, step e % PT_C).
    // The line, its base index and skew of each element are fixed for the tile.
    long ebase[EL];
    int eskew[EL];
    bool eok[EL], eshared[EL];
#pragma unroll
    for (int m = 0; m < EL; ++m) {
        const int e = tid + NT * m;
        const int L = e / PT_C;
        line_info(L, &ebase[m], &eskew[m], &eok[m]);
        const int a = L % PT_J, b = L / PT_J;
        const int ad = FWD ? a : PT_J - 1 - a, bd = FWD ? b : PT_K - 1 - b;
        eshared[m] = ad == PT_J - 1 || bd == PT_K - 1;
        eskew[m] -= e % PT_C;  // fold the step offset u into the skew
    }
    T lr[EL], lv[EL];
    uint32_t lc[EL];
    T lej = 0, lejv = 0;  // one edge element per lane (lanes < PT_K: -j edge, next PT_J lanes: -k edge)
    auto elem_cell = [&](int m, int s0, long* c) {
        const int il = s0 - eskew[m];
        const int i = FWD ? il : g.ni - 1 - il;
        *c = ebase[m] + i;
        return eok[m] && il >= 0 && il < g.ni;
    };
    auto issue_loads = [&](int s0) {
#pragma unroll
        for (int m = 0; m < EL; ++m) {
            long c;
            const bool ok = elem_cell(m, s0, &c);
            lr[m] = ok ? in[c] : (T)0;
            lv[m] = ok ? inv[c] : (T)0;
            lc[m] = ok ? (uint32_t)code[c] : 0u;
        }
        // edge lines: the neighbour's cells next to our jd = 0 (kd = 0) lanes
        lej = 0; lejv = 0;
        if (tid < PT_K * PT_C && hasJ) {
            const int b = tid / PT_C, u = tid % PT_C;      // line (jd = 0, kl = b)
            const int a = FWD ? 0 : PT_J - 1;
            long bc; int sk; bool ok; int i;
            line_info(b * PT_J + a, &bc, &sk, &ok);
            ok = ok && cell_of(s0 + u, sk, &i);
            if (ok) {
                lej = load_shared<T, FENCE>(&out[bc + i + offJ]);
                lejv = inv[bc + i + offJ];
            }
        } else if (tid >= 128 && tid < 128 + PT_J * PT_C && hasK) {
            const int t2 = tid - 128;
            const int a = t2 / PT_C, u = t2 % PT_C;         // line (jl = a, kd = 0)
            const int b = FWD ? 0 : PT_K - 1;
            long bc; int sk; bool ok; int i;
            line_info(b * PT_J + a, &bc, &sk, &ok);
            ok = ok && cell_of(s0 + u, sk, &i);
            if (ok) {
                lej = load_shared<T, FENCE>(&out[bc + i + offK]);
                lejv = inv[bc + i + offK];
            }
        }
    };
    auto commit_loads = [&](int buf) {
#pragma unroll
        for (int m = 0; m < EL; ++m) {
            const int e = tid + NT * m;
            const int L = e / PT_C, u = e % PT_C;
            sr[buf][u][L] = lr[m];
            si[buf][u][L] = lv[m];
            sc[buf][u][L] = lc[m];
        }
        if (tid < PT_K * PT_C) {
            ejq[buf][tid % PT_C][tid / PT_C] = lej;
            eji[buf][tid % PT_C][tid / PT_C] = lejv;
        } else if (tid >= 128 && tid < 128 + PT_J * PT_C) {
            const int t2 = tid - 128;
            ekq[buf][t2 % PT_C][t2 / PT_C] = lej;
            eki[buf][t2 % PT_C][t2 / PT_C] = lejv;
        }
    };
    auto flush_chunk = [&](int s0) {
#pragma unroll
        for (int m = 0; m < EL; ++m) {
            const int e = tid + NT * m;
            const int L = e / PT_C, u = e % PT_C;
            const int s = s0 + u;
            long c;
            if (s < 0 || s >= total || !elem_cell(m, s0, &c)) continue;
            const T v = lq[s % (2 * PT_C)][L / PT_J][L % PT_J];
            if (eshared[m]) store_shared<T, FENCE>(&out[c], v);
            else out[c] = v;
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
    // edge lane (jd = 0) at step s uses the neighbour's step s + PT_J - 1
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
            if constexpr (FENCE) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    };

    if (tid == 0) { knownJ = 0; knownK = 0; }
    __syncthreads();
    wait_chunk(0);
    __syncthreads();
    issue_loads(0);
    commit_loads(0);

    T prevq = 0, previ = 0;
    for (int t = 0; t < nchunks; ++t) {
        const int s0 = t * PT_C, s1 = s0 + PT_C, cur = t & 1;
        // (1)-(2): everything issued before this point is complete -> the
        // flush of chunk t-2 (issued at the top of chunk t-1) is visible
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __syncthreads();
        if (tid == 0 && t >= 2) {
            if constexpr (FENCE) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(&prog[myTile], base + (unsigned long long)(s0 - PT_C), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        // (3)
        if (s1 < total) wait_chunk(s1);
        __syncthreads();
        // (4) next chunk's inputs, (5) previous chunk's outputs
        if (s1 < total) issue_loads(s1);
        if (t >= 1) flush_chunk(s0 - PT_C);
#pragma unroll
        for (int u = 0; u < PT_C; ++u) {
            const int s = s0 + u;
            if (s < total) {
                const int pb = (s + 1) & 1, cb = s & 1;
                const int rp = (s + 2 * PT_C - 1) % (2 * PT_C), rc = s % (2 * PT_C);
                const int il = s - skew;
                if (valid && il >= 0 && il < g.ni) {
                    const uint16_t cd = (uint16_t)sc[cur][u][tid];
                    T v = sr[cur][u][tid];
                    const T ic = si[cur][u][tid];
                    // each update is computed unconditionally and kept only when the
                    // entry is present, which leaves the serial arithmetic unchanged
                    const int ck = code_dir(cd, dK), cj = code_dir(cd, dJ), ci2 = code_dir(cd, dI);
                    const T ak = (T)code_coef(ck), aj = (T)code_coef(cj), ai = (T)code_coef(ci2);
                    if (FWD) {
                        // column order c-nn, c-ni, c-1; L_{c,m} = A_{c,m} inv_m
                        const T qk = kd > 0 ? lq[rp][kl - 1][jl] : ekq[cur][u][jl];
                        const T ik = kd > 0 ? li[pb][kl - 1][jl] : eki[cur][u][jl];
                        const T qj = jd > 0 ? lq[rp][kl][jl - 1] : ejq[cur][u][kl];
                        const T ij = jd > 0 ? li[pb][kl][jl - 1] : eji[cur][u][kl];
                        T w = v - (ak * ik) * qk;
                        v = present_cls(ck) ? w : v;
                        w = v - (aj * ij) * qj;
                        v = present_cls(cj) ? w : v;
                        w = v - (ai * previ) * prevq;
                        v = present_cls(ci2) ? w : v;
                    } else {
                        // row order c+1, c+ni, c+nn; L_{m,c} = A_{m,c} inv_c
                        const T qj = jd > 0 ? lq[rp][kl][jl + 1] : ejq[cur][u][kl];
                        const T qk = kd > 0 ? lq[rp][kl + 1][jl] : ekq[cur][u][jl];
                        T w = v - (ai * ic) * prevq;
                        v = present_cls(ci2) ? w : v;
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
        if (s1 < total) commit_loads(cur ^ 1);
    }
    // flush the last chunk and publish completion
    __syncthreads();
    flush_chunk((nchunks - 1) * PT_C);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (tid == 0) {
        if constexpr (FENCE) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&prog[myTile], base + (unsigned long long)total, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
}

// One workgroup per tile, tiles in topological (diagonal) block order.
template <class T, bool FWD, int PT_J, int PT_K, int PT_C, bool FENCE>
__global__ __launch_bounds__(PT_J * PT_K) void k_sweep_pipe(Geo g, int ntj, int ntk, const uint16_t* __restrict__ code,
                                                            const T* __restrict__ inv, const T* __restrict__ in,
                                                            T* __restrict__ out, const PcgScalars* S,
                                                            unsigned long long* prog, unsigned long long epoch,
                                                            int* err) {
    if (S && S->done) return;
    int J, K;
    tile_of_block(blockIdx.x, ntj, ntk, FWD, &J, &K);
    sweep_tile<T, FWD, PT_J, PT_K, PT_C, FENCE>(J, K, g, ntj, ntk, code, inv, in, out, prog, epoch, err);
}

// Single-XCD schedule: the grid holds 8 workgroups per tile; only those on
// XCD 0 (read from the hardware register, not assumed) work, claiming tiles in
// topological order from a ticket counter until none are left.
template <class T, bool FWD, int PT_J, int PT_K, int PT_C, bool FENCE>
__global__ __launch_bounds__(PT_J * PT_K) void k_sweep_pipe_xcd(Geo g, int ntj, int ntk,
                                                                const uint16_t* __restrict__ code,
                                                                const T* __restrict__ inv, const T* __restrict__ in,
                                                                T* __restrict__ out, const PcgScalars* S,
                                                                unsigned long long* prog, unsigned long long epoch,
                                                                int* err, int* ticket) {
    __shared__ int my;
    if (S && S->done) return;
    const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u;  // HW_REG_XCC_ID
    if (xcc != 0) return;
    const int tiles = ntj * ntk;
    for (;;) {
        if (threadIdx.x == 0) my = atomicAdd(ticket, 1);
        __syncthreads();
        const int t = my;
        __syncthreads();
        if (t >= tiles) return;
        int J, K;
        tile_of_block(t, ntj, ntk, FWD, &J, &K);
        sweep_tile<T, FWD, PT_J, PT_K, PT_C, FENCE>(J, K, g, ntj, ntk, code, inv, in, out, prog, epoch, err);
    }
}

// Schedules (sweep_kernel), all 16x16 tiles:
//   2: sc1 hand-off, 8-step chunks   3: sc1, 4-step chunks
//   4: fenced hand-off, 4-step       5: fenced, single XCD, 4-step
//   6: sc1, single XCD, 4-step
// Flags layout (ints): [0] error word, [1] pad, [2, 2+2T) 64-bit progress
// counters (T tiles of 16x16), then a ring of 256 ticket counters.
template <class T, int C, bool FENCE>
static void launch_pipe(const Geo& g, bool forward, bool xcd, const uint16_t* code, const T* inv, const T* in,
                        T* out, const PcgScalars* S, unsigned long long* prog, unsigned long long ep, int* err,
                        int* ticket, hipStream_t st) {
    constexpr int TJ = 16, TK = 16;
    const int ntj = (g.nj + TJ - 1) / TJ, ntk = (g.nk + TK - 1) / TK;
    const int tiles = ntj * ntk;
    if (xcd) {
        if (forward)
            k_sweep_pipe_xcd<T, true, TJ, TK, C, FENCE><<<8 * tiles, TJ * TK, 0, st>>>(g, ntj, ntk, code, inv, in,
                                                                                      out, S, prog, ep, err, ticket);
        else
            k_sweep_pipe_xcd<T, false, TJ, TK, C, FENCE><<<8 * tiles, TJ * TK, 0, st>>>(g, ntj, ntk, code, inv, in,
                                                                                       out, S, prog, ep, err, ticket);
    } else {
        if (forward)
            k_sweep_pipe<T, true, TJ, TK, C, FENCE><<<tiles, TJ * TK, 0, st>>>(g, ntj, ntk, code, inv, in, out, S,
                                                                              prog, ep, err);
        else
            k_sweep_pipe<T, false, TJ, TK, C, FENCE><<<tiles, TJ * TK, 0, st>>>(g, ntj, ntk, code, inv, in, out, S,
                                                                               prog, ep, err);
    }
}

static int tiles16(const Geo& g) { return ((g.nj + 15) / 16) * ((g.nk + 15) / 16); }

template <class T>
bool sweep_pipelined(const Geo& g, int kind, bool forward, const uint16_t* code, const T* inv, const T* in, T* out,
                     const PcgScalars* S, int* flags, unsigned long long* epoch, hipStream_t st) {
    const int tiles = tiles16(g);
    if (!flags || !epoch || tiles > 4096) return false;
    unsigned long long* prog = reinterpret_cast<unsigned long long*>(flags + 2);
    int* ring = flags + 2 + 2 * tiles;
    const unsigned long long ep = (*epoch)++;
    if (ep % 256 == 0) hipMemsetAsync(ring, 0, 256 * sizeof(int), st);
    int* ticket = ring + (ep % 256);
    switch (kind) {
        case 3: launch_pipe<T, 4, false>(g, forward, false, code, inv, in, out, S, prog, ep, flags, ticket, st); break;
        case 4: launch_pipe<T, 4, true>(g, forward, false, code, inv, in, out, S, prog, ep, flags, ticket, st); break;
        case 5: launch_pipe<T, 4, true>(g, forward, true, code, inv, in, out, S, prog, ep, flags, ticket, st); break;
        case 6: launch_pipe<T, 4, false>(g, forward, true, code, inv, in, out, S, prog, ep, flags, ticket, st); break;
        default: launch_pipe<T, 8, false>(g, forward, false, code, inv, in, out, S, prog, ep, flags, ticket, st); break;
    }
    return true;
}
int sweep_flag_words(const Geo& g) { return 2 + 2 * tiles16(g) + 256; }
template bool sweep_pipelined<double>(const Geo&, int, bool, const uint16_t*, const double*, const double*, double*,
                                      const PcgScalars*, int*, unsigned long long*, hipStream_t);
template bool sweep_pipelined<float>(const Geo&, int, bool, const uint16_t*, const float*, const float*, float*,
                                     const PcgScalars*, int*, unsigned long long*, hipStream_t);
