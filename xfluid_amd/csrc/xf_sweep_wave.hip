// xf_sweep_wave.hip -- the MIC(0) triangular sweeps (pcg_solver.h:192-220) and
// the MIC(0) factorisation (pcg_solver.h:85-186) as pipelined line-tile
// schedules, one wavefront per 8x8 tile of i-lines.
//
// A compute wavefront owns a tile: lane jl + 8*kl holds line (jl, kl) and walks
// the tile's skewed hyperplanes (forward step s: the cells with i + jl + kl = s;
// backward: the same slots in reverse order).  At step s the -i neighbour of a
// cell is the same register one step earlier, the -j neighbour lane-1's register
// (a DPP row shift), the -k neighbour lane-8's (ds_bpermute).  A step needs no
// barrier and no LDS round trip for the recurrence itself.  Backward sweeps
// mirror every direction.
//
// The per-cell coefficients L_cm = A_cm * inv_m (forward: the neighbour's
// pivot, pcg_solver.h:198-201; backward: the own pivot, :214-216) are formed
// once per solve ("records").  A missing entry keeps the running value bit for
// bit (the serial sweep does not visit it), so each term is evaluated as
// w = v - L*q and kept through a bitwise select for absent entries.
//
// Tiles pass edge values to their downstream neighbours as tagged words: a
// value is stored together with a 32-bit tag in one 8-byte word (FP64: two
// words, one per half), with device-scope stores.  A reader that sees the tag
// has the value, so no progress counter and no wait for store completion is
// involved.  Waits are bounded: on timeout an error bit is set in flags[0] and
// the kernel still runs to completion, and the host redoes the solve with
// per-hyperplane sweeps.
#include <hip/hip_ext.h>

#include "xf_internal.h"

#define WV_SPIN (1 << 20)       // LDS polls before a wait inside the workgroup is declared stuck
#define WV_SPIN_EDGE (1 << 16)  // re-reads of an edge word (each a device-scope round trip)

typedef __attribute__((address_space(3))) unsigned char lds_u8;
typedef __attribute__((address_space(3))) int lds_int;

// error word of the sweeps (flags[0]) and a spare
int sweep_flag_words(const Geo&) { return 2; }

// tiles by diagonal J+K, upstream tiles first (a tile only ever waits on tiles
// that were dispatched before it)
__device__ inline void wv_tile(int b, int ntj, int ntk, bool fwd, int* J, int* K) {
    int d = 0;
    for (;; ++d) {
        const int j0 = d - (ntk - 1) > 0 ? d - (ntk - 1) : 0;
        const int j1 = d < ntj - 1 ? d : ntj - 1;
        const int cnt = j1 - j0 + 1;
        if (b < cnt) { *J = j0 + b; *K = d - *J; break; }
        b -= cnt;
    }
    if (!fwd) { *J = ntj - 1 - *J; *K = ntk - 1 - *K; }
}

// Tile of workgroup b in the fused kernels, whose tiles are all resident at
// once, so any one-to-one map is valid.  Workgroups are dispatched to the 8
// XCDs in turn (b % 8); each XCD gets a compact block of tiles (bands of 4
// k-rows walked j by j, cut into 8 consecutive runs), so most tile-to-tile
// hand-offs stay within an XCD, whose device-coherent round trips are shorter.
__device__ inline void w8_tile_xcd(int b, int ntj, int ntk, int* J, int* K) {
    const int tiles = ntj * ntk, q = tiles >> 3, r = tiles & 7, x = b & 7, m = b >> 3;
    const int p = x < r ? x * (q + 1) + m : r * (q + 1) + (x - r) * q + m;
    constexpr int H = 4;
    const int band = p / (H * ntj), k0 = band * H, hb = ntk - k0 < H ? ntk - k0 : H;
    const int w = p - band * H * ntj;
    *J = w / hb;
    *K = k0 + w % hb;
}

// Sum over the 64 lanes of a wave (every lane gets it), in a fixed order: the
// backward sweep's dot(z, r) partials and k_dot_tiles share it, so both give the
// same bits.
__device__ inline double w8_wave_sum(double v) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) v = v + __shfl_xor(v, o);
    return v;
}

template <int N>
__device__ inline void vm_wait() {
    static_assert(N >= 0 && N < 64, "vmcnt field");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// LDS accesses of the load waves.  The compiler orders any LDS access after an
// LDS-DMA load behind a full vmcnt wait (it cannot tell the addresses apart),
// which would drain the loads in flight at every step; these accesses are
// written out instead, each waiting for its own completion.
__device__ inline int lds_ld32(uint32_t a) {
    int v;
    asm volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(a) : "memory");
    return v;
}
__device__ inline uint64_t lds_ld64(uint32_t a) {
    uint64_t v;
    asm volatile("ds_read_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(v) : "v"(a) : "memory");
    return v;
}
__device__ inline void lds_st32(uint32_t a, int v) {
    asm volatile("ds_write_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" ::"v"(a), "v"(v) : "memory");
}
// store without waiting (LDS operations of a wave complete in order, so an
// extra one in flight only makes the compiler's own lgkmcnt waits stricter)
__device__ inline void lds_st32_nowait(uint32_t a, int v) {
    asm volatile("ds_write_b32 %0, %1" ::"v"(a), "v"(v) : "memory");
}
__device__ inline void lds_st64(uint32_t a, uint64_t v) {
    asm volatile("ds_write_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" ::"v"(a), "v"(v) : "memory");
}

// lane-1 (forward) / lane+1 (backward) inside each 16-lane row
template <bool FWD>
__device__ inline int wv_dpp(int x) {
    return __builtin_amdgcn_update_dpp(0, x, FWD ? 0x111 : 0x101, 0xf, 0xf, true);
}
template <bool FWD>
__device__ inline float wv_shift(float x) { return __int_as_float(wv_dpp<FWD>(__float_as_int(x))); }
template <bool FWD>
__device__ inline double wv_shift(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = wv_dpp<FWD>((int)b), hi = wv_dpp<FWD>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (uint32_t)lo);
}
__device__ inline float wv_perm(int a, float x) { return __int_as_float(__builtin_amdgcn_ds_bpermute(a, __float_as_int(x))); }
__device__ inline double wv_perm(int a, double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_bpermute(a, (int)b), hi = __builtin_amdgcn_ds_bpermute(a, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (uint32_t)lo);
}

typedef uint32_t wv_u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t wv_u32x2 __attribute__((ext_vector_type(2)));
#define WV_SC1 16  // cache policy: device-coherent access

// ---------------------------------------------------------------------------
// 8x8 tiles.  A sweep step moves every input of a tile through one CU, and one
// CU fetches only some 20-25 GB/s from memory, so the step of a 16x16 tile is
// bound by its bytes rather than its arithmetic.  Tiles of 8x8 lines (one cell
// per lane) put four times as many CUs on the same grid with a quarter of the
// bytes per step each.  A step's records arrive LEAD steps ahead of the
// compute wave through a record wave; an edge wave follows the upstream edge
// words one step ahead (every step of lead would be paid again on each
// tile-to-tile hop).  The two are separate waves because a wave's vector-memory
// counter completes in issue order: waiting for an edge would also wait for
// every record load issued before it.
//
// Oct layout: element (slot, lane) of tile t = J*ntk + K at (t*SL + slot)*64 +
// lane, lane = jl + 8*kl, SL = ni + 14; cell (i, 8J + jl, 8K + kl) with
// i = slot - jl - kl.  Records: four values per element, {L0, L1, L2, pivot}
// (forward -k, -j, -i; backward +i, +j, +k).  An absent entry is stored as +0
// (its term then leaves the sum's value as it is) and a present one as A_cm * inv.  Edge rows: 16 words per tile and step, 0-7 the edge column
// by kl, 8-15 the edge row by jl, whose corner cell is not repeated there
// (readers take it from the column word).  An edge word's tag is the sweep's
// tag base plus the producing step (w8_next_tag), so a ring slot's copy of an
// older step, or a word of an earlier launch, never passes for the current one.
#define W8_PAD 14
#define W8_B 16  // steps per unrolled block of the compute wave


// Geometry of the schedule on one of two axes.  The lines (a cell per lane and
// step) run along axis L; the lanes cover axis A (lane & 7: the -1 neighbour is
// one lane down) and axis B (lane >> 3: eight lanes down).  ax 0: L = i, A = j,
// B = k; ax 2: L = k, A = i, B = j.  Tile (J, K) covers A in [8J, 8J + 8) and B in
// [8K, 8K + 8); its index is J * ntB + K.  Strides are the natural layout's.
struct W8Geo {
    int ax, nL, nA, nB, ntA, ntB, SL;
    long sL, sA, sB, N;
};
__host__ __device__ inline W8Geo w8_geo(const Geo& g, int ax) {
    W8Geo G;
    G.ax = ax;
    if (ax == 2) {
        G.nL = g.nk; G.nA = g.ni; G.nB = g.nj;
        G.sL = g.nn; G.sA = 1; G.sB = g.ni;
    } else {
        G.nL = g.ni; G.nA = g.nj; G.nB = g.nk;
        G.sL = 1; G.sA = g.ni; G.sB = g.nn;
    }
    G.ntA = (G.nA + 7) / 8;
    G.ntB = (G.nB + 7) / 8;
    G.SL = G.nL + W8_PAD;
    G.N = g.N;
    return G;
}
// edge rows per tile: one per step and 7 written after the last step (+1 spare)
__host__ __device__ inline int w8_erows(const W8Geo& G) { return G.SL + 8; }
// buffer sizes: the larger of the two axes
long sweep_w8_size(const Geo& g) {
    long n = 0;
    for (int ax : {0, 2}) {
        const W8Geo G = w8_geo(g, ax);
        const long m = (long)G.ntA * G.ntB * G.SL * 64;
        n = m > n ? m : n;
    }
    return n;
}
long sweep_w8_edge_words(const Geo& g) {
    long n = 0;
    for (int ax : {0, 2}) {
        const W8Geo G = w8_geo(g, ax);
        const long m = (long)G.ntA * G.ntB * w8_erows(G) * 16 * 4;  // u32 units, sized for FP64 words
        n = m > n ? m : n;
    }
    return n;
}

// natural-layout cell (i, j, k) of element e; false if it lies off the grid
__device__ inline bool w8_cell(const W8Geo& G, long e, int* i, int* j, int* k) {
    const int lane = (int)(e & 63);
    const long row = e >> 6;
    const long t = row / G.SL;
    const int slot = (int)(row - t * G.SL);
    const int jl = lane & 7, kl = lane >> 3;
    const int a = (int)(t / G.ntB) * 8 + jl, b = (int)(t % G.ntB) * 8 + kl, l = slot - jl - kl;
    if (G.ax == 2) { *i = a; *j = b; *k = l; }
    else { *i = l; *j = a; *k = b; }
    return a < G.nA && b < G.nB && l >= 0 && l < G.nL;
}

struct Box8 { int i0, i1, j0, j1, k0, k1; };
// cells outside the active box are zero (the sweep did not visit their slots)
// The elements the sweeps over a box read: its tiles' slots [s0, s0 + nS).
// Thread t of a launch over ntjb * ntkb * nS * 64 threads -> element e.
struct W8BoxArgs { int J0, K0, ntkb, s0, nS, ntjb; };
__device__ inline bool w8_box_element(const W8Geo& G, const W8BoxArgs& b, long t, long* e) {
    const long total = (long)b.ntjb * b.ntkb * b.nS * 64;
    if (t >= total) return false;
    const int lane = (int)(t & 63);
    const long r = t >> 6;
    const int sr = (int)(r % b.nS), tb = (int)(r / b.nS);
    const long tile = (long)(b.J0 + tb / b.ntkb) * G.ntB + (b.K0 + tb % b.ntkb);
    *e = (tile * G.SL + b.s0 + sr) * 64 + lane;
    return true;
}

template <class T>
__global__ void k_from_oct(Geo g, W8Geo G, const T* __restrict__ src, T* __restrict__ dst, const PcgScalars* S,
                           Box8 b) {
    if (S && S->done) return;
    const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const int i = (int)(c % g.ni), j = (int)((c / g.ni) % g.nj), k = (int)(c / g.nn);
    if (i < b.i0 || i > b.i1 || j < b.j0 || j > b.j1 || k < b.k0 || k > b.k1) { dst[c] = (T)0; return; }
    const int l = G.ax == 2 ? k : i, a = G.ax == 2 ? i : j, bb = G.ax == 2 ? j : k;
    const long t = (long)(a >> 3) * G.ntB + (bb >> 3);
    const int jl = a & 7, kl = bb & 7;
    dst[c] = src[(t * G.SL + l + jl + kl) * 64 + jl + 8 * kl];
}

// dot(z, r) partials of the tiles of the box in the backward sweep's order
// (per lane over the slots from last to first, then w8_wave_sum), from z and r
// in the natural layout: the per-hyperplane sweeps give the same partials as
// the pipelined ones.  One 64-lane block per tile.
template <class T>
__global__ __launch_bounds__(64) void k_dot_tiles(W8Geo G, int J0, int K0, int ntkb, int s0, int nS,
                                                  const T* __restrict__ z, const T* __restrict__ r,
                                                  double* __restrict__ partial, const PcgScalars* S) {
    if (S && S->done) return;
    const int tb = blockIdx.x, J = J0 + tb / ntkb, K = K0 + tb % ntkb;
    const int jl = threadIdx.x & 7, kl = threadIdx.x >> 3, a = 8 * J + jl, b = 8 * K + kl;
    double acc = 0.0;
    if (a < G.nA && b < G.nB) {
        const long base = G.sA * a + G.sB * b;
        for (int t = 0; t < nS; ++t) {
            const int l = s0 + nS - 1 - t - jl - kl;
            if ((unsigned)l < (unsigned)G.nL) {
                const long c = base + G.sL * l;
                acc = acc + (double)z[c] * (double)r[c];
            }
        }
    }
    acc = w8_wave_sum(acc);
    if (threadIdx.x == 0) partial[tb] = acc;
}

// records: rec[4e..4e+3] forward, rec[4n + 4e..] backward
template <class T>
__global__ void k_w8_records(Geo g, W8Geo G, long n, W8BoxArgs bx, const uint16_t* __restrict__ code,
                             const T* __restrict__ inv, T* __restrict__ rec) {
    long e;
    if (!w8_box_element(G, bx, blockIdx.x * (long)blockDim.x + threadIdx.x, &e)) return;
    int i, j, k;
    T lf[3] = {0, 0, 0}, lb[3] = {0, 0, 0}, ic = 0;
    if (w8_cell(G, e, &i, &j, &k)) {
        const long c = cidx(g, i, j, k);
        const uint16_t cd = code[c];
        ic = inv[c];
        const int dF[3] = {D_ZM, D_YM, D_XM}, dB[3] = {D_XP, D_YP, D_ZP};
        const long oF[3] = {-g.nn, -(long)g.ni, -1};
        const bool inF[3] = {k > 0, j > 0, i > 0};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int cf = code_dir(cd, dF[d]), cb = code_dir(cd, dB[d]);
            // A_cm < 0 and inv >= 0: the product carries the sign bit (-0 included)
            if ((cf == 1 || cf == 2) && inF[d]) lf[d] = (T)code_coef(cf) * inv[c + oF[d]];
            if (cb == 1 || cb == 2) lb[d] = (T)code_coef(cb) * ic;
        }
    }
    HIP_vector_type<T, 2>* f = (HIP_vector_type<T, 2>*)(rec + 4 * e);
    HIP_vector_type<T, 2>* b = (HIP_vector_type<T, 2>*)(rec + 4 * n + 4 * e);
    f[0] = HIP_vector_type<T, 2>(lf[0], lf[1]);
    f[1] = HIP_vector_type<T, 2>(lf[2], ic);
    b[0] = HIP_vector_type<T, 2>(lb[0], lb[1]);
    b[1] = HIP_vector_type<T, 2>(lb[2], ic);
}

template <class T>
struct W8Planes {
    const T* in;     // input: natural layout (forward sweep), (q, r) pairs in the oct layout (backward)
    T* out;          // forward output: (q, r) pairs in the oct layout (the pivots: q alone)
    const T* rec;    // the sweep's records
    uint32_t* edge;  // tagged edge words
    T* nat;          // backward output z, natural layout
    double* dot;     // backward: per-tile partials of dot(z, r), or null
};


// edge-row word holding the upstream's value for lane jl of the edge row
template <bool FWD>
__device__ inline int w8_kword(int jl) {
    constexpr int CJ = FWD ? 7 : 0;
    return jl == CJ ? CJ : 8 + jl;
}
// wait until ring slot s is free (step s - R consumed); cons = last count seen.
// POLL: also give up once another wave has set the error word.  The record
// wave does not poll it: a load with a register result in its loop, even on a
// path it rarely takes, lets the compiler put a full vmcnt wait before its
// publishing store, which would drain the record requests in flight (the
// compute wave it waits on keeps publishing after any failed wait of its own,
// so the spin limit alone bounds this wait).
template <int R, bool POLL = true>
__device__ inline void w8_wait_slot(int s, int& cons, uint32_t consA, int* err) {
    if (s < R) return;
    int spins = 0;
    while (cons < s - R + 1) {
        cons = lds_ld32(consA);
        if (cons >= s - R + 1) break;
        __builtin_amdgcn_s_sleep(1);
        ++spins;
        if (POLL && (spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) cons = s;
        else if (spins > WV_SPIN) { atomicOr(err, 4); cons = s; }
    }
}

// Runs the tiles [J0, J0 + ntjb) x [K0, K0 + ntkb) of the ntk-wide tile grid over
// the slots [s0, s0 + nS): the tiles and slots that hold active rows (all other
// rows are empty and their results zero).  Step t works on slot s0 + t forward,
// s0 + nS - 1 - t backward; edge rows are indexed by step.  After its last step
// a tile writes 7 more edge rows of zeros (steps nS .. nS + 6): a downstream
// tile's last 7 steps wait for those, for cells outside the box whose entries
// are absent, so every step checks its edge words the same way.
//
// The ring holds R steps (32; 16 for the pivots, whose records are twice as
// long; 64 for the fused FP32 sweep when every tile has a CU of its own, DEEP,
// 99 KB of LDS, one workgroup per CU), the records of LEAD = 8 steps in
// flight.  The record wave issues a step's requests in about 0.1 us, so the
// first step's records arrive some 0.4 us + 0.1 us x LEAD after the wave
// starts (device-clock traces: 1.2 us at 8, 3.5 us at 31 steps in flight),
// while 8 steps in flight keep the records ahead of a compute wave stepping
// at 0.16 us (fewer in flight measured the same step time).  The compute wave
// runs in blocks of W8_B steps unrolled so that every ring slot sits at a
// fixed LDS offset from the block's base.  A step is bound by latency (an LDS
// exchange and the dependent multiply-add chain), not by issue: eight more
// independent instructions per step left its time unchanged.  Lanes without an
// upstream edge read the slot's dummy word (value 0, the step's tag), which
// makes the tag check two compares for every lane; the record count is checked
// every 4 steps and the consumed count published every 4 steps (which needs
// R - LEAD >= 4 slots of slack).  The record wave's waits are the explicit
// vm_wait counts: a compiler-inserted full vmcnt wait before its publishing
// LDS store (a register allocation that reuses a load's registers can cause
// one) would drain every step's requests and double the step time.
//
// MIC: the same pipeline computes the MIC(0) pivots (factor_modified_incomplete_
// cholesky0, pcg_solver.h:85-186, in the order of k_mic_plane): the value
// exchanged between neighbours is the pivot, the records are {cL0, cL1, cL2,
// diag, cM0, cM1, cM2, 0} with cL the entry's coefficient (+0 when absent: with
// finite neighbour pivots both updates then leave d unchanged bit for bit) and
// cM the sum of the neighbour's coefficients that the fill-in drops, and
// there is no input vector.
// LDS ring of the 8x8-tile pipeline: depth in steps and bytes per step
template <class T, bool MIC, bool DEEP = false>
struct W8Ring {
    static_assert(!DEEP || (sizeof(T) == 4 && !MIC), "deep ring: FP32 sweeps");
    static constexpr int NS = sizeof(T) / 4;
    static constexpr int R = MIC ? 16 : DEEP ? 64 : 32;
    static constexpr int SLOT = 64 * (MIC ? 8 : 4) * (int)sizeof(T) + (MIC ? 0 : 128 * (int)sizeof(T)) + 2 * 16 * 8 * NS + 16;
};

// One sweep of tile (J, K) by the workgroup's three waves (ring_ and ctr_ are
// the workgroup's LDS).
// PRE: the record wave's loads of steps 0 .. LEAD - 1 were issued already (by
// w8_prefetch with the same arguments), so its loop starts at step 0.
template <class T, bool FWD, bool MIC, bool DEEP, int AX, bool PRE = false>
__device__ __attribute__((always_inline)) inline void w8_sweep(const W8Geo& G, int J, int K, int J0, int K0, int ntjb,
                                                               int ntkb, int s0, int nS, W8Planes<T> P,
                                                               uint32_t tagbase, int* err, double tau, double sigma,
                                                               unsigned char* ring_, int* ctr_) {
    static_assert(!MIC || (FWD && sizeof(T) == 8), "pivots: forward, FP64");
    constexpr int NS = sizeof(T) / 4;
    constexpr int R = W8Ring<T, MIC, DEEP>::R;                      // ring depth in steps
    constexpr int LEAD = 8;  // steps of record loads in flight
    static_assert((R & (R - 1)) == 0 && R % W8_B == 0 && LEAD + 4 <= R, "ring depth");
    constexpr int EWB = 8 * NS;                  // bytes per tagged edge word
    constexpr int ROWB = 16 * EWB;               // bytes per edge row
    constexpr int RECB = 64 * (MIC ? 8 : 4) * sizeof(T);  // record bytes per step (1, 2 or 4 KB)
    constexpr int INB = MIC ? 0 : 128 * sizeof(T);        // input bytes per step (room for the backward pairs)
    constexpr int EOFF = RECB + INB;             // upstream edge rows (j, then k) within a slot
    constexpr int DOFF = EOFF + 2 * ROWB;        // the slot's dummy word
    constexpr int EB = 2 * ROWB + 16;            // edge bytes per slot
    constexpr int SLOT = EOFF + EB;
    // the forward sweep gathers its input from the natural layout (one 4-byte
    // load per lane and word: no separate layout pass) and stores (q, r) pairs in
    // the oct layout; the backward sweep reads those pairs (16-byte loads), so it
    // has r for dot(z, r) at hand, and stores z in the natural layout
    constexpr bool NATIN = FWD && !MIC;
    constexpr bool PAIR = FWD && !MIC;
    constexpr int PERD = RECB / 1024 + (INB ? (NATIN ? NS : 1) : 0);  // record + input loads per step
    constexpr int CJ = FWD ? 7 : 0;              // jl of the edge column (and of the row's corner)
    static_assert(SLOT == W8Ring<T, MIC, DEEP>::SLOT, "ring layout");
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int SL = G.SL, ER = w8_erows(G);
    const long myTile = (long)J * G.ntB + K;
    const int uJ = FWD ? J - 1 : J + 1, uK = FWD ? K - 1 : K + 1;
    const bool hasJ = uJ >= J0 && uJ < J0 + ntjb, hasK = uK >= K0 && uK < K0 + ntkb;
    auto slot_of = [s0, nS](int t) __attribute__((always_inline)) { return FWD ? s0 + t : s0 + nS - 1 - t; };
    const long tJ = (long)uJ * G.ntB + K, tK = (long)J * G.ntB + uK;
    lds_u8* ring = (lds_u8*)ring_;
    lds_int* ctr = (lds_int*)ctr_;
    const uint32_t ctrA = (uint32_t)(uintptr_t)ctr, ringA = (uint32_t)(uintptr_t)ring;
    if (threadIdx.x < 2) ctr[threadIdx.x] = 0;
    // the edge and dummy words of every slot start out invalid (tags are never zero)
    for (int b = threadIdx.x * 4; b < R * EB; b += 192 * 4)
        *(__attribute__((address_space(3))) uint32_t*)(ring + (b / EB) * SLOT + EOFF + b % EB) = 0u;
    __syncthreads();

    if (wid == 1) {
        // ------------------------------- record wave -------------------------------
        // at iteration x: step x waited for and published, then the records and
        // input of step x + LEAD issued (LDS-DMA into the ring).  Publishing
        // first keeps the count independent of the slot that step x + LEAD
        // waits for: the compute wave asks for up to 5 steps ahead of the last
        // count it published, with R - LEAD >= 4 slots of slack.
        int cons_seen = 0;
        const int rjl = lane & 7, rkl = lane >> 3;
        const bool rjk = 8 * J + rjl < G.nA && 8 * K + rkl < G.nB;
        // cell of slot t: rrow + sL * t (the line's cell l = t - rjl - rkl); the
        // step's part sL * t is wave-uniform, so no per-lane multiply
        const T* rrow = P.in + (rjk ? G.sA * (8 * J + rjl) + G.sB * (8 * K + rkl) - G.sL * (rjl + rkl) : 0);
        const int rl0 = rjl + rkl;
        for (int x = PRE ? 0 : -LEAD; x < nS; ++x) {
            if (x >= 0) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                // (steps x + 1 .. x + LEAD - 1 may stay in flight)
                if (x + LEAD - 1 < nS) vm_wait<(LEAD - 1) * PERD>();
                else vm_wait<0>();
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                lds_st32(ctrA, x + 1);
            }
            const int s = x + LEAD;
            if (s < nS) {
                w8_wait_slot<R, false>(s, cons_seen, ctrA + 4, err);
                const long e0 = (myTile * SL + slot_of(s)) * 64;
                lds_u8* dst = ring + (s % R) * SLOT;
#pragma unroll
                for (int h = 0; h < RECB / 1024; ++h)
                    __builtin_amdgcn_global_load_lds((const unsigned char*)(P.rec + (MIC ? 8 : 4) * e0) + h * 1024 +
                                                         lane * 16,
                                                     dst + h * 1024, 16, 0, 0);
                if constexpr (NATIN) {
                    // lane's cell at this slot, or cell 0 for a lane outside the grid
                    // (its records are zero, so its input only has to be finite)
                    const int t = slot_of(s);
                    const bool ok = rjk & ((unsigned)(t - rl0) < (unsigned)G.nL);
                    const long so = G.sL * t;
                    const unsigned char* a = (const unsigned char*)(ok ? rrow + so : P.in);
                    __builtin_amdgcn_global_load_lds(a, dst + RECB, 4, 0, 0);
                    if (NS == 2) __builtin_amdgcn_global_load_lds(a + 4, dst + RECB + 256, 4, 0, 0);
                } else if (INB && lane < INB / 16) {
                    __builtin_amdgcn_global_load_lds((const unsigned char*)(P.in + 2 * e0) + lane * 16, dst + RECB, 16, 0,
                                                     0);
                }
            }
        }
        // (the last step reads "next" records from slot nS, which nothing uses)
        lds_st32(ctrA, nS + 1);
        return;
    }
    if (wid == 2) {
        // -------------------------------- edge wave --------------------------------
        // Lanes 0-15 each follow one upstream edge word (0-7: the j-upstream's
        // column word kl, 8-15: the k-upstream's row word for jl): the word of
        // step x + 2 is loaded while step x is checked; a stale word (older tag)
        // is re-read at once until it carries the tag of its step, then copied
        // into the ring slot, where the compute wave checks the tag once more.
        // Lane 16 writes the slot's dummy word.  (With the words of two steps in
        // flight the wave takes three steps per memory round trip; one step ahead
        // capped a tile that waits on its upstream at two steps per round trip,
        // about 0.2 us a step; three ahead return stale more often: traced
        // slower.)
        const bool mine = lane < 16, sideJ = lane < 8, dummy = lane == 16;
        const bool has = mine && (sideJ ? hasJ : hasK);
        const long ut = has ? (sideJ ? tJ : tK) : myTile;
        const int w = mine ? (sideJ ? lane : w8_kword<FWD>(lane - 8)) : 0;
        const int off = dummy ? DOFF : EOFF + (sideJ ? 0 : ROWB) + w * EWB;
        // the upstream's word of its step x + 7 (its rows run to step nS + 6)
        auto src = [&](int x) __attribute__((always_inline)) {
            return (const uint64_t*)((const unsigned char*)P.edge + (ut * ER + x + 7) * ROWB + w * EWB);
        };
        int cons_seen = 0;
        // one step: load the words of step x + 2 into (n0, n1) while checking and
        // placing those of step x (c0, c1); three register sets rotate, so no
        // copy makes the compiler wait for a load just issued
        auto step = [&](int x, uint64_t& c0, uint64_t& c1, uint64_t& n0, uint64_t& n1) __attribute__((always_inline)) {
            const int xn = x + 2 < nS ? x + 2 : nS - 1;
            n0 = __hip_atomic_load(src(xn), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (NS == 2) n1 = __hip_atomic_load(src(xn) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t want = tagbase + (uint32_t)(x + 7);
            if (has && x < nS) {
                auto fresh = [&](uint64_t v0, uint64_t v1) __attribute__((always_inline)) {
                    return (uint32_t)(v0 >> 32) == want && (NS == 1 || (uint32_t)(v1 >> 32) == want);
                };
                if (!fresh(c0, c1)) {
                    // stale: two re-reads in flight, the second issued about half a
                    // round trip after the first, each re-issued as soon as it returns
                    // stale, so a word is seen within about half a trip of landing
                    int spins = 0;
                    uint64_t e0 = 0, e1 = 0;
                    c0 = __hip_atomic_load(src(x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (NS == 2) c1 = __hip_atomic_load(src(x) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __builtin_amdgcn_s_sleep(4);
                    e0 = __hip_atomic_load(src(x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (NS == 2) e1 = __hip_atomic_load(src(x) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    for (;;) {
                        if (fresh(c0, c1)) break;
                        if ((spins & 63) == 63 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                        if (++spins > WV_SPIN_EDGE) { atomicOr(err, 2); break; }
                        c0 = __hip_atomic_load(src(x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (NS == 2) c1 = __hip_atomic_load(src(x) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (fresh(e0, e1)) { c0 = e0; c1 = e1; break; }
                        e0 = __hip_atomic_load(src(x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (NS == 2) e1 = __hip_atomic_load(src(x) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
            w8_wait_slot<R>(x, cons_seen, ctrA + 4, err);
            if ((mine || dummy) && x < nS) {
                const uint64_t d = (uint64_t)want << 32;
                const uint32_t ea = ringA + (x % R) * SLOT + off;
                asm volatile("ds_write_b64 %0, %1" ::"v"(ea), "v"(dummy ? d : c0) : "memory");
                if (NS == 2) asm volatile("ds_write_b64 %0, %1 offset:8" ::"v"(ea), "v"(dummy ? d : c1) : "memory");
            }
        };
        uint64_t a0 = 0, a1 = 0, b0 = 0, b1 = 0, c0 = 0, c1 = 0;
        const int x1 = 1 < nS ? 1 : nS - 1;
        a0 = __hip_atomic_load(src(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (NS == 2) a1 = __hip_atomic_load(src(0) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b0 = __hip_atomic_load(src(x1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (NS == 2) b1 = __hip_atomic_load(src(x1) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // (three steps per trip keep the register sets in place at the loop head;
        // a step x >= nS only issues a harmless load of step nS - 1's word)
        for (int x = 0; x < nS; x += 3) {
            step(x, a0, a1, c0, c1);
            step(x + 1, b0, b1, a0, a1);
            step(x + 2, c0, c1, b0, b1);
        }
        return;
    }

    // ------------------------------- compute wave -------------------------------
    const int jl = lane & 7, kl = lane >> 3;
    const bool eJ = FWD ? jl == 0 : jl == 7;  // j-neighbour in the upstream tile
    const bool eK = FWD ? kl == 0 : kl == 7;  // k-neighbour in the upstream tile
    const bool oJ = jl == CJ;                 // edge column (read by the j-downstream tile)
    const bool oK = kl == (FWD ? 7 : 0);      // edge row (read by the k-downstream tile)
    const int bp = ((lane + (FWD ? 56 : 8)) & 63) * 4;  // lane -8 / +8 (edge lanes are overridden)
    // each lane reads two edge words per step: its upstream words where it has
    // them, the dummy word otherwise
    const int ejo = eJ && hasJ ? EOFF + kl * EWB : DOFF;
    const int eko = eK && hasK ? EOFF + ROWB + w8_kword<FWD>(jl) * EWB : DOFF;
    const int eout = oJ ? kl * EWB : (8 + jl) * EWB;
    // a lane off both edges stores its words out of the edge buffer's range
    // (buffer stores beyond num_records are dropped by the hardware)
    const int estore = oJ || oK ? eout : 0x7ffff000;
    // this tile's output plane and edge rows (offsets in bytes from the tile's start)
    constexpr int OPC = PAIR ? 2 : 1;  // values per cell in the output plane
    const __amdgpu_buffer_rsrc_t orc = __builtin_amdgcn_make_buffer_rsrc(
        FWD ? P.out + myTile * SL * 64 * OPC : nullptr, (short)0, FWD ? SL * 64 * OPC * (int)sizeof(T) : 0, 0x00020000);
    // backward: z in the natural layout, from the tile's first cell; a lane off the
    // grid (or a cell with i outside [0, ni)) stores out of the buffer's range
    const long norg = G.sA * 8 * J + G.sB * 8 * K;
    const bool njk = 8 * J + jl < G.nA && 8 * K + kl < G.nB;
    // byte offset of step s's cell: zb - s * sLb (a wave-uniform product), its line
    // cell l = ls - s
    const int sLb = (int)G.sL * (int)sizeof(T), ls = s0 + nS - 1 - jl - kl;
    const int zb = (int)(G.sA * jl + G.sB * kl) * (int)sizeof(T) + ls * sLb;
    const long nrem = FWD || !P.nat ? 0 : (G.N - norg) * (long)sizeof(T);
    const __amdgpu_buffer_rsrc_t zrc = __builtin_amdgcn_make_buffer_rsrc(
        FWD || !P.nat ? nullptr : P.nat + norg, (short)0, (int)(nrem < 0x7fffff00L ? nrem : 0x7fffff00L), 0x00020000);
    double dacc = 0.0;  // backward: this lane's sum of z * r
    const __amdgpu_buffer_rsrc_t erc = __builtin_amdgcn_make_buffer_rsrc(
        (unsigned char*)P.edge + myTile * ER * ROWB, (short)0, ER * ROWB, 0x00020000);
    typedef T T2v __attribute__((ext_vector_type(2)));
    typedef T T4v __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) const uint64_t ldsW;
    typedef __attribute__((address_space(3))) const wv_u32x4 ldsW4;
    typedef __attribute__((address_space(3))) const T ldsT;
    struct Rec { T L0, L1, L2, ic, rin, M0, M1, M2, rd; };
    int seen = 0;
    // wait until the record wave has published `upto` steps
    auto need_rec = [&](int upto) __attribute__((always_inline)) {
        if (seen < upto) {
            int spins = 0;
            for (;;) {
                seen = __builtin_amdgcn_readfirstlane(lds_ld32(ctrA));
                if (seen >= upto) break;
                __builtin_amdgcn_s_sleep(1);
                // (another wave's failed wait ends this one too)
                if ((++spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    seen = nS + 1;
                    break;
                }
                if (spins > 4 * WV_SPIN) { atomicOr(err, 8); seen = nS + 1; break; }
            }
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    };
    // records + input of the step in ring slot `slot`
    auto load_rec = [&](int slot, Rec& R) __attribute__((always_inline)) {
        const lds_u8* sl = ring + slot * SLOT;
        if constexpr (MIC) {
            // (diag in the ic field)
            const T2v a = *(__attribute__((address_space(3))) const T2v*)(sl + lane * 64);
            const T2v b = *(__attribute__((address_space(3))) const T2v*)(sl + lane * 64 + 16);
            const T2v c = *(__attribute__((address_space(3))) const T2v*)(sl + lane * 64 + 32);
            const T2v d = *(__attribute__((address_space(3))) const T2v*)(sl + lane * 64 + 48);
            R.L0 = a.x; R.L1 = a.y; R.L2 = b.x; R.ic = b.y; R.M0 = c.x; R.M1 = c.y; R.M2 = d.x; R.rin = 0;
        } else if constexpr (NS == 1) {
            const T4v r4 = *(__attribute__((address_space(3))) const T4v*)(sl + lane * 16);
            R.L0 = r4.x; R.L1 = r4.y; R.L2 = r4.z; R.ic = r4.w;
            if constexpr (NATIN) {
                R.rin = *(ldsT*)(sl + RECB + lane * sizeof(T));
            } else {
                const T2v pr = *(__attribute__((address_space(3))) const T2v*)(sl + RECB + lane * 2 * sizeof(T));
                R.rin = pr.x; R.rd = pr.y;
            }
        } else {
            const T2v a = *(__attribute__((address_space(3))) const T2v*)(sl + lane * 32);
            const T2v b = *(__attribute__((address_space(3))) const T2v*)(sl + lane * 32 + 16);
            R.L0 = a.x; R.L1 = a.y; R.L2 = b.x; R.ic = b.y;
            if constexpr (NATIN) {
                typedef __attribute__((address_space(3))) const uint32_t ldsU;
                const uint32_t lo = *(ldsU*)(sl + RECB + lane * 4), hi = *(ldsU*)(sl + RECB + 256 + lane * 4);
                R.rin = (T)__longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
            } else {
                const T2v pr = *(__attribute__((address_space(3))) const T2v*)(sl + RECB + lane * 2 * sizeof(T));
                R.rin = pr.x; R.rd = pr.y;
            }
        }
    };
    // an edge word as read from the ring: FP32 {value, tag}, FP64 {lo, tag, hi, tag}
    auto ew_bad = [](wv_u32x4 w, uint32_t want) __attribute__((always_inline)) {
        return NS == 1 ? w.y != want : (w.y != want) | (w.w != want);
    };
    auto ew_val = [](wv_u32x4 w) __attribute__((always_inline)) {
        if constexpr (NS == 1) return (T)__uint_as_float(w.x);
        else return (T)__longlong_as_double((long long)(((uint64_t)w.z << 32) | w.x));
    };
    auto ew_read = [&](const lds_u8* a) __attribute__((always_inline)) {
        wv_u32x4 w;
        if constexpr (NS == 1) {
            const uint64_t v = *(ldsW*)a;
            w = wv_u32x4{(uint32_t)v, (uint32_t)(v >> 32), 0u, 0u};
        } else {
            w = *(ldsW4*)a;
        }
        return w;
    };
    T qp = 0;
    // One step.  The edge words, the neighbour exchange of the previous step's
    // results and the next step's records are all requested before the first
    // wait, so their latencies overlap; then the edge tags are checked and the
    // dependent chain runs on records read one step earlier.
    // (s = sb + u, u <= ub: the store offsets split into a per-block part and
    // a constant one that the instructions carry)
    auto step = [&](int s, int slot, Rec& cur, Rec& nxt, int sb, int u, int ub) __attribute__((always_inline)) {
        const lds_u8* sl = ring + slot * SLOT;
        const uint32_t want = tagbase + (uint32_t)(s + 7);
        wv_u32x4 wj = ew_read(sl + ejo), wk = ew_read(sl + eko);
        __builtin_amdgcn_sched_barrier(0);
        T qj = wv_shift<FWD>(qp);
        T qk = wv_perm(bp, qp);
        __builtin_amdgcn_sched_barrier(0);
        load_rec((slot + 1) & (R - 1), nxt);
        __builtin_amdgcn_sched_barrier(0);
        if (__builtin_amdgcn_ballot_w64(ew_bad(wj, want) | ew_bad(wk, want))) {
            // not yet placed by the edge wave: re-read (bounded) until the tags match,
            // with no sleep between the reads (a word placed just after a read is
            // seen one LDS round trip later: 1 % faster sweeps than sleeping first)
            int spins = 0;
            const uint32_t aj = ringA + slot * SLOT + ejo, ak = ringA + slot * SLOT + eko;
            do {
                if ((++spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (spins > 4 * WV_SPIN) { atomicOr(err, 16); break; }
                // (each 8-byte read holds a value half with its own tag)
                const uint64_t j0 = lds_ld64(aj), k0 = lds_ld64(ak);
                const uint64_t j1 = NS == 2 ? lds_ld64(aj + 8) : 0, k1 = NS == 2 ? lds_ld64(ak + 8) : 0;
                wj = wv_u32x4{(uint32_t)j0, (uint32_t)(j0 >> 32), (uint32_t)j1, (uint32_t)(j1 >> 32)};
                wk = wv_u32x4{(uint32_t)k0, (uint32_t)(k0 >> 32), (uint32_t)k1, (uint32_t)(k1 >> 32)};
            } while (__builtin_amdgcn_ballot_w64(ew_bad(wj, want) | ew_bad(wk, want)));
        }
        qj = eJ ? ew_val(wj) : qj;
        qk = eK ? ew_val(wk) : qk;
        T v = cur.rin, q;
        // the neighbour values in the records' direction order (forward -k, -j, -i,
        // backward +i, +j, +k): qp is the line's previous cell, qj the A neighbour,
        // qk the B neighbour
        const T n0 = FWD ? (AX == 2 ? qp : qk) : (AX == 2 ? qj : qp);
        const T n1 = FWD ? (AX == 2 ? qk : qj) : (AX == 2 ? qk : qj);
        const T n2 = FWD ? (AX == 2 ? qj : qp) : (AX == 2 ? qp : qk);
        if constexpr (MIC) {
            // d = diag - sum over the -k, -j, -i entries of L^2 + tau L miss
            const T diag = cur.ic;
            T d = diag;
            T L = cur.L0 * n0, miss = cur.M0 * n0;
            d -= L * L;
            d -= (T)tau * L * miss;
            L = cur.L1 * n1; miss = cur.M1 * n1;
            d -= L * L;
            d -= (T)tau * L * miss;
            L = cur.L2 * n2; miss = cur.M2 * n2;
            d -= L * L;
            d -= (T)tau * L * miss;
            q = diag == 0 ? (T)0 : (d < (T)sigma * diag) ? 1 / sqrt(diag) : 1 / sqrt(d);
        } else if constexpr (NS == 1) {
            // FP32: fused multiply-adds, the line's previous cell first and the B
            // neighbour (permuted across rows, the last to arrive) last;
            // k_fwd_plane / k_bwd_plane<float> use the same order, so the fallback
            // gives the same bits.  (Record index of the own, A and B entries.)
            // An absent entry is stored as +0 and every neighbour value here is
            // finite (zeros outside the grid and the box), so its term leaves v's
            // value as it is: no test for presence.
            constexpr int IO = FWD ? (AX == 2 ? 0 : 2) : (AX == 2 ? 2 : 0);
            constexpr int IA = FWD ? (AX == 2 ? 2 : 1) : (AX == 2 ? 0 : 1);
            constexpr int IB = 3 - IO - IA;
            const T Ls[3] = {cur.L0, cur.L1, cur.L2};
            v = __builtin_fmaf(-Ls[IO], qp, v);
            v = __builtin_fmaf(-Ls[IA], qj, v);
            v = __builtin_fmaf(-Ls[IB], qk, v);
            q = v * cur.ic;
        } else {
            // FP64: the reference's order (forward c-nn, c-ni, c-1; backward c+1,
            // c+ni, c+nn), each term rounded as there; an absent entry (+0, with a
            // finite neighbour value) leaves v's value as it is
            v = v - cur.L0 * n0;
            v = v - cur.L1 * n1;
            v = v - cur.L2 * n2;
            q = v * cur.ic;
        }
        const uint64_t qb = NS == 1 ? (uint64_t)__float_as_uint((float)q) : (uint64_t)__double_as_longlong((double)q);
        // edge words first (the downstream tiles wait for them; the output store can
        // wait): every lane stores, those off the edge out of the buffer's range
        const uint32_t t = tagbase + (uint32_t)s;
        if constexpr (NS == 1) {
            __builtin_amdgcn_raw_buffer_store_b64(wv_u32x2{(uint32_t)qb, t}, erc, estore + u * ROWB, sb * ROWB, WV_SC1);
        } else {
            __builtin_amdgcn_raw_buffer_store_b128(wv_u32x4{(uint32_t)qb, t, (uint32_t)(qb >> 32), t}, erc,
                                                   estore + u * ROWB, sb * ROWB, WV_SC1);
        }
        if constexpr (FWD) {
            // (q, r) pairs (the pivots: q), step by step in the tile's plane
            constexpr int SB = 64 * OPC * (int)sizeof(T);
            const int so = (s0 + sb) * SB, oi = u * SB;
            if constexpr (!PAIR) {
                __builtin_amdgcn_raw_buffer_store_b64(wv_u32x2{(uint32_t)qb, (uint32_t)(qb >> 32)}, orc, lane * 8 + oi,
                                                      so, 0);
            } else if constexpr (NS == 1) {
                __builtin_amdgcn_raw_buffer_store_b64(wv_u32x2{(uint32_t)qb, __float_as_uint((float)cur.rin)}, orc,
                                                      lane * 8 + oi, so, 0);
            } else {
                const uint64_t rb = (uint64_t)__double_as_longlong((double)cur.rin);
                __builtin_amdgcn_raw_buffer_store_b128(
                    wv_u32x4{(uint32_t)qb, (uint32_t)(qb >> 32), (uint32_t)rb, (uint32_t)(rb >> 32)}, orc, lane * 16 + oi,
                    so, 0);
            }
        } else {
            // z at cell l of this lane's line; dot(z, r) in the step order
            const bool ok = njk & ((unsigned)(ls - s) < (unsigned)G.nL);
            const int zo = ok ? zb - s * sLb : 0x7fffff80;
            if constexpr (NS == 1) __builtin_amdgcn_raw_buffer_store_b32((uint32_t)qb, zrc, zo, 0, 0);
            else __builtin_amdgcn_raw_buffer_store_b64(wv_u32x2{(uint32_t)qb, (uint32_t)(qb >> 32)}, zrc, zo, 0, 0);
            dacc = dacc + (double)q * (double)cur.rd;
        }
        qp = q;
    };
    // steps consumed: their slots may be refilled (their reads are issued, and a
    // wave's LDS operations complete in order)
    auto publish = [&](int n) __attribute__((always_inline)) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        lds_st32_nowait(ctrA + 4, n);
    };
    Rec ra, rb;
    need_rec(1);
    load_rec(0, ra);
    // (settle every load issued so far, kernel arguments included, so the step
    // loop's waits only ever concern its own requests)
    __builtin_amdgcn_s_waitcnt(0);
    int s = 0;
    // two record sets alternate (no register copies between steps)
    for (; s + W8_B <= nS; s += W8_B) {
        const int h = s & (R - 1);  // the block's first slot
#pragma unroll
        for (int u = 0; u < W8_B; u += 4) {
            need_rec(s + u + 5 < nS + 1 ? s + u + 5 : nS + 1);
            step(s + u, h + u, ra, rb, s, u, W8_B - 1);
            step(s + u + 1, h + u + 1, rb, ra, s, u + 1, W8_B - 1);
            step(s + u + 2, h + u + 2, ra, rb, s, u + 2, W8_B - 1);
            step(s + u + 3, h + u + 3, rb, ra, s, u + 3, W8_B - 1);
            publish(s + u + 4);
        }
    }
    for (; s < nS; s += 2) {
        need_rec(s + 3 < nS + 1 ? s + 3 : nS + 1);
        step(s, s & (R - 1), ra, rb, s, 0, 0);
        if (s + 1 < nS) step(s + 1, (s + 1) & (R - 1), rb, ra, s + 1, 0, 0);
        publish(s + 2 < nS ? s + 2 : nS);
    }
    if constexpr (!FWD) {
        // the tile's partial of dot(z, r) (k_dot_tiles sums in the same order)
        dacc = w8_wave_sum(dacc);
        if (P.dot && lane == 0) P.dot[(J - J0) * ntkb + (K - K0)] = dacc;
    }
    // edge rows of the steps after the last: zeros, for the downstream tiles' last 7 steps
    if (oJ || oK) {
        for (int d = 0; d < 7; ++d) {
            const uint32_t t = tagbase + (uint32_t)(nS + d);
            if constexpr (NS == 1) {
                __builtin_amdgcn_raw_buffer_store_b64(wv_u32x2{0u, t}, erc, eout, (nS + d) * ROWB, WV_SC1);
            } else {
                __builtin_amdgcn_raw_buffer_store_b128(wv_u32x4{0u, t, 0u, t}, erc, eout, (nS + d) * ROWB, WV_SC1);
            }
        }
    }
}

template <class T, bool FWD, bool MIC, int AX>
__global__ __launch_bounds__(192) void k_sweep_w8(W8Geo G, int J0, int K0, int ntjb, int ntkb, int s0, int nS,
                                                  W8Planes<T> P, const PcgScalars* S, uint32_t tagbase, int* err,
                                                  double tau, double sigma) {
    __shared__ __attribute__((aligned(16))) unsigned char ring_[W8Ring<T, MIC>::R * W8Ring<T, MIC>::SLOT];
    __shared__ int ctr_[2];  // [0] steps whose records are in, [1] steps consumed
    if (S && S->done) return;
    int J, K;
    wv_tile(blockIdx.x, ntjb, ntkb, FWD, &J, &K);
    w8_sweep<T, FWD, MIC, false, AX>(G, J + J0, K + K0, J0, K0, ntjb, ntkb, s0, nS, P, tagbase, err, tau, sigma, ring_,
                                     ctr_);
}

// The forward sweep's record wave requests for steps 0 .. LEAD - 1 (records and
// natural-layout input into their ring slots), as its loop in w8_sweep issues
// them; a kernel issues these before its own prologue and then runs
// w8_sweep<..., PRE = true>.  Nothing the sweep computes is read.
template <class T, bool DEEP>
__device__ __attribute__((always_inline)) inline void w8_prefetch(const W8Geo& G, int J, int K, int s0, int nS,
                                                                  const W8Planes<T>& P, unsigned char* ring_) {
    constexpr int NS = sizeof(T) / 4, LEAD = 8, RECB = 64 * 4 * sizeof(T);
    constexpr int R = W8Ring<T, false, DEEP>::R, SLOT = W8Ring<T, false, DEEP>::SLOT;
    if ((threadIdx.x >> 6) != 1) return;
    const int lane = threadIdx.x & 63, rjl = lane & 7, rkl = lane >> 3, rl0 = rjl + rkl;
    const bool rjk = 8 * J + rjl < G.nA && 8 * K + rkl < G.nB;
    const T* rrow = P.in + (rjk ? G.sA * (8 * J + rjl) + G.sB * (8 * K + rkl) - G.sL * (rjl + rkl) : 0);
    const long myTile = (long)J * G.ntB + K;
    lds_u8* ring = (lds_u8*)ring_;
    for (int st = 0; st < LEAD && st < nS; ++st) {
        const int t = s0 + st;
        const long e0 = (myTile * G.SL + t) * 64;
        lds_u8* dst = ring + (st % R) * SLOT;
#pragma unroll
        for (int h = 0; h < RECB / 1024; ++h)
            __builtin_amdgcn_global_load_lds((const unsigned char*)(P.rec + 4 * e0) + h * 1024 + lane * 16,
                                             dst + h * 1024, 16, 0, 0);
        const bool ok = rjk & ((unsigned)(t - rl0) < (unsigned)G.nL);
        const unsigned char* a = (const unsigned char*)(ok ? rrow + G.sL * t : P.in);
        __builtin_amdgcn_global_load_lds(a, dst + RECB, 4, 0, 0);
        if (NS == 2) __builtin_amdgcn_global_load_lds(a + 4, dst + RECB + 256, 4, 0, 0);
    }
}

// Forward and backward sweep in one launch: each workgroup sweeps its tile
// forward, then the same tile backward (the tile that ends the forward sweep
// begins the backward one, with no launch in between).  Every tile waits on
// tiles that start later, so all tiles must be resident at once: the host
// uses this kernel only when they fit (w8_fused_fits).
template <class T, bool DEEP, int AX>
__global__ __launch_bounds__(192) void k_sweep_w8_fb(W8Geo G, int J0, int K0, int ntjb, int ntkb, int s0,
                                                     int nS, W8Planes<T> Pf, W8Planes<T> Pb, const PcgScalars* S,
                                                     uint32_t tag1, uint32_t tag2, int* err, const double* fin_partial,
                                                     int fin_n, int fin_it, PcgScalars* Sw, PcgScalars* mirror) {
    __shared__ __attribute__((aligned(16))) unsigned char ring_[W8Ring<T, false, DEEP>::R * W8Ring<T, false, DEEP>::SLOT];
    __shared__ int ctr_[2];
    int J, K;
    w8_tile_xcd(blockIdx.x, ntjb, ntkb, &J, &K);
    // the forward sweep's first records and input are requested before the
    // residual test, which they do not depend on: their latency overlaps its own
    w8_prefetch<T, DEEP>(G, J + J0, K + K0, s0, nS, Pf, ring_);
    if (fin_partial) {
        // the iteration's residual test (k_finish FIN_RES): every workgroup reduces the
        // partials in the same order and agrees on `stop`; the workgroup of the tile
        // that starts last (it waits on every other tile anyway) records it
        __shared__ double sh[16];
        __shared__ int stop;
        // all of a thread's loads in flight at once, the solver state's among them
        // (the partials were just written by another kernel, so each load is a
        // round trip to memory)
        constexpr int FU = 6;
        double v[FU];
#pragma unroll
        for (int u = 0; u < FU; ++u) {
            const int q = threadIdx.x + u * 192;
            v[u] = q < fin_n ? fin_partial[q] : 0.0;
        }
        int was = 0;
        double tol = 0.0;
        if (threadIdx.x == 0) {
            was = Sw->done;
            tol = Sw->tol;
        }
        double m = 0.0;
#pragma unroll
        for (int u = 0; u < FU; ++u) m = v[u] > m ? v[u] : m;
        for (int q = threadIdx.x + FU * 192; q < fin_n; q += 192) m = fin_partial[q] > m ? fin_partial[q] : m;
        m = block_reduce<double, true>(m, sh);
        if (threadIdx.x == 0) {
            stop = was != 0 || m <= tol;
            if (J == ntjb - 1 && K == ntkb - 1) {
                if (!was) {
                    Sw->res = m;
                    Sw->it = fin_it + 1;
                    if (m <= tol) { Sw->done = 1; Sw->iters = fin_it + 1; }
                }
                if (mirror) *mirror = *Sw;
            }
        }
        __syncthreads();
        if (stop) {
            vm_wait<0>();  // (the requests land before the workgroup ends)
            return;
        }
    } else if (S && S->done) {
        vm_wait<0>();
        return;
    }
    w8_sweep<T, true, false, DEEP, AX, true>(G, J + J0, K + K0, J0, K0, ntjb, ntkb, s0, nS, Pf, tag1, err, 0.0, 0.0, ring_,
                                   ctr_);
    // this tile's forward output is its backward input: stores complete, LDS free
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    w8_sweep<T, false, false, DEEP, AX>(G, J + J0, K + K0, J0, K0, ntjb, ntkb, s0, nS, Pb, tag2, err, 0.0, 0.0, ring_,
                                    ctr_);
}

// MIC(0) records (k_sweep_w8<double, true, true>): {cL0, cL1, cL2, diag, cM0, cM1,
// cM2, 0} per element, directions -k, -j, -i; cM sums the neighbour's
// coefficients in the two directions k_mic_plane drops (exact halves, so
// cM * inv rounds like the sum of the two products).
__global__ void k_mic_records(Geo g, W8Geo G, W8BoxArgs bx, const uint16_t* __restrict__ code,
                              double* __restrict__ rec) {
    long e;
    if (!w8_box_element(G, bx, blockIdx.x * (long)blockDim.x + threadIdx.x, &e)) return;
    int i, j, k;
    double cL[3] = {0, 0, 0}, cM[3] = {0, 0, 0}, diag = 0;
    if (w8_cell(G, e, &i, &j, &k)) {
        const long c = cidx(g, i, j, k);
        const uint16_t cc = code[c];
        diag = code_diag(cc);
        const int dF[3] = {D_ZM, D_YM, D_XM};
        const int ma[3] = {D_XP, D_XP, D_YP}, mb[3] = {D_YP, D_ZP, D_ZP};
        const long oF[3] = {-g.nn, -(long)g.ni, -1};
        const bool inF[3] = {k > 0, j > 0, i > 0};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int cl = code_dir(cc, dF[d]);
            if ((cl == 1 || cl == 2) && inF[d]) {
                cL[d] = code_coef(cl);
                const uint16_t nc = code[c + oF[d]];
                const int a = code_dir(nc, ma[d]), b = code_dir(nc, mb[d]);
                double m = 0.0;
                if (a == 1 || a == 2) m += code_coef(a);
                if (b == 1 || b == 2) m += code_coef(b);
                cM[d] = m;
            }
        }
    }
    double* r = rec + 8 * e;
    r[0] = cL[0]; r[1] = cL[1]; r[2] = cL[2]; r[3] = diag;
    r[4] = cM[0]; r[5] = cM[1]; r[6] = cM[2]; r[7] = 0.0;
}

__global__ void k_w8_clear(long n, uint32_t* __restrict__ p) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e < n) p[e] = 0u;
}

// Tag base of the next launch on this edge buffer: bit 31 set, a 15-bit launch
// count in bits 16-30 and the step in the low 16 bits.  When the count wraps
// the buffer is cleared, so no word left from 32768 launches earlier can carry
// a current tag.
template <class T>
static bool w8_next_tag(const Geo& g, const W8Geo& G, const SweepSkew<T>& sk, uint32_t* tag, hipStream_t st) {
    if (w8_erows(G) >= 0xffff) return false;
    const uint32_t n = sk.w8_launch++ & 0x7fffu;
    if (n == 0 && sk.w8_launch > 1) {
        const long words = sweep_w8_edge_words(g);
        k_w8_clear<<<xf_blocks(words), XF_BLOCK, 0, st>>>(words, sk.oEdge);
    }
    *tag = 0x80000000u | (n << 16);
    return true;
}

// the box's tiles and slots on axis G.ax: tiles [J0, J0 + ntjb) x [K0, K0 + ntkb),
// slots [s0, s0 + nS); box = (i0, i1, j0, j1, k0, k1)
struct W8Box { int J0, K0, ntjb, ntkb, s0, nS; };
static W8Box w8_box(const W8Geo& G, const int* box) {
    const int* L = G.ax == 2 ? box + 4 : box;      // line axis range
    const int* A = G.ax == 2 ? box : box + 2;      // A range
    const int* B = G.ax == 2 ? box + 2 : box + 4;  // B range
    W8Box w;
    w.J0 = A[0] / 8;
    w.K0 = B[0] / 8;
    w.ntjb = A[1] / 8 - w.J0 + 1;
    w.ntkb = B[1] / 8 - w.K0 + 1;
    w.s0 = L[0];
    const int s1 = L[1] + W8_PAD < G.SL ? L[1] + W8_PAD : G.SL - 1;
    w.nS = s1 - w.s0 + 1;
    return w;
}

// Axis of the lines for a box: each tile-to-tile hand-off on the critical path
// costs about as much as 20 steps, so the lines run along the axis that leaves
// the fewest tiles across the box's two other extents (ties: along i).  Lines
// along k need natural-layout offsets of a whole grid within 31 bits.
int w8_choose_axis(const Geo& g, const int* box) {
    long best = -1;
    int ax = 0;
    for (int a : {0, 2}) {
        if (a == 2 && g.N * 8 >= 0x7fff0000L) continue;
        const W8Box w = w8_box(w8_geo(g, a), box);
        const long cost = 20L * (w.ntjb + w.ntkb - 2) + w.nS;
        if (best < 0 || cost < best) { best = cost; ax = a; }
    }
    return ax;
}

bool mic0_w8(const Geo& g, const uint16_t* code, double* inv, double tau, double sigma, const SweepSkew<double>& sk,
             int* flags, unsigned long long* epoch, hipStream_t st) {
    const W8Geo G = w8_geo(g, sk.ax);
    const long n = sweep_w8_size(g);
    const long ebytes = (long)G.ntA * G.ntB * w8_erows(G) * 16 * 16;
    uint32_t tag;
    if (!flags || !sk.oRec || ebytes >= 0x7fffffffL || !w8_next_tag(g, G, sk, &tag, st)) return false;
    const W8Box w = w8_box(G, sk.box);
    const W8BoxArgs ba{w.J0, w.K0, w.ntkb, w.s0, w.nS, w.ntjb};
    k_mic_records<<<xf_blocks((long)w.ntjb * w.ntkb * w.nS * 64), XF_BLOCK, 0, st>>>(g, G, ba, code, sk.oRec);
    const W8Planes<double> P{nullptr, sk.oQ, sk.oRec, sk.oEdge, nullptr, nullptr};
    auto k = sk.ax == 2 ? k_sweep_w8<double, true, true, 2> : k_sweep_w8<double, true, true, 0>;
    k<<<w.ntjb * w.ntkb, 192, 0, st>>>(G, w.J0, w.K0, w.ntjb, w.ntkb, w.s0, w.nS, P, nullptr, tag, flags, tau, sigma);
    const Box8 bx{sk.box[0], sk.box[1], sk.box[2], sk.box[3], sk.box[4], sk.box[5]};
    k_from_oct<double><<<xf_blocks(g.N), XF_BLOCK, 0, st>>>(g, G, sk.oQ, inv, nullptr, bx);
    return true;
}

template <class T>
void sweep_prepare_w8(const Geo& g, const uint16_t* code, const T* inv, const SweepSkew<T>& sk, hipStream_t st) {
    // (only the elements of the box's tiles and slots: the sweeps read no others)
    const long n = sweep_w8_size(g);
    const W8Geo G = w8_geo(g, sk.ax);
    const W8Box w = w8_box(G, sk.box);
    const W8BoxArgs ba{w.J0, w.K0, w.ntkb, w.s0, w.nS, w.ntjb};
    k_w8_records<T><<<xf_blocks((long)w.ntjb * w.ntkb * w.nS * 64), XF_BLOCK, 0, st>>>(g, G, n, ba, code, inv,
                                                                                       sk.oRec);
}

// whether all `tiles` workgroups of k_sweep_w8_fb<T, DEEP, *> can be resident at once
template <class T, bool DEEP>
static bool w8_fused_fits(int tiles) {
    static int cap = -1;
    if (cap < 0) {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_sweep_w8_fb<T, DEEP, 0>, 192, 0) != hipSuccess)
            cus = per = 0;
        cap = cus * per;
    }
    return tiles <= cap;
}

int w8_box_tiles(const Geo& g, int ax, const int* box) {
    const W8Box w = w8_box(w8_geo(g, ax), box);
    return w.ntjb * w.ntkb;
}

template <class T>
void w8_dot_tiles(const Geo& g, int ax, const int* box, const T* z, const T* r, double* partial, const PcgScalars* S,
                  hipStream_t st) {
    const W8Geo G = w8_geo(g, ax);
    const W8Box w = w8_box(G, box);
    k_dot_tiles<T><<<w.ntjb * w.ntkb, 64, 0, st>>>(G, w.J0, w.K0, w.ntkb, w.s0, w.nS, z, r, partial, S);
}

template <class T, int AX>
static void w8_launch(const W8Geo& G, const W8Box& w, bool fused, bool deep, const W8Planes<T>& f,
                      const W8Planes<T>& b, const PcgScalars* S, uint32_t tag1, uint32_t tag2, int* flags,
                      FinRes* fin, hipEvent_t t0, hipEvent_t t1, hipStream_t st) {
    const int tiles = w.ntjb * w.ntkb;
    // t0 / t1 ride on the dispatch packets themselves, so they time the kernels and not
    // the marker packets (and their cache release) that hipEventRecord would add
    if (fused) {
        const double* fp = fin ? fin->partial : nullptr;
        hipExtLaunchKernelGGL(deep ? k_sweep_w8_fb<T, sizeof(T) == 4, AX> : k_sweep_w8_fb<T, false, AX>, dim3(tiles),
                              dim3(192), 0, st, t0, t1, 0, G, w.J0, w.K0, w.ntjb, w.ntkb, w.s0, w.nS, f, b, S, tag1,
                              tag2, flags, fp, fin ? fin->n : 0, fin ? fin->it : 0,
                              fin ? fin->S : (PcgScalars*)nullptr, fin ? fin->mirror : (PcgScalars*)nullptr);
        if (fin) fin->ran = true;
    } else {
        if (fin) {
            launch_finish_res(*fin, st);
            fin->ran = true;
        }
        hipExtLaunchKernelGGL(k_sweep_w8<T, true, false, AX>, dim3(tiles), dim3(192), 0, st, t0, (hipEvent_t) nullptr,
                              0, G, w.J0, w.K0, w.ntjb, w.ntkb, w.s0, w.nS, f, S, tag1, flags, 0.0, 0.0);
        hipExtLaunchKernelGGL(k_sweep_w8<T, false, false, AX>, dim3(tiles), dim3(192), 0, st, (hipEvent_t) nullptr, t1,
                              0, G, w.J0, w.K0, w.ntjb, w.ntkb, w.s0, w.nS, b, S, tag2, flags, 0.0, 0.0);
    }
}

template <class T>
bool sweep_precond_w8(const Geo& g, const SweepSkew<T>& sk, const T* r, T* z, const PcgScalars* S, int* flags,
                      unsigned long long* epoch, hipEvent_t t0, hipEvent_t t1, hipStream_t st, double* dot_partial,
                      FinRes* fin) {
    const W8Geo G = w8_geo(g, sk.ax);
    const long n = sweep_w8_size(g);
    const long ebytes = (long)G.ntA * G.ntB * w8_erows(G) * 16 * 8 * (long)(sizeof(T) / 4);
    uint32_t tag1, tag2;
    if (!flags || !sk.oRec || ebytes >= 0x7fffffffL || !w8_next_tag(g, G, sk, &tag1, st) ||
        !w8_next_tag(g, G, sk, &tag2, st))
        return false;
    const W8Box w = w8_box(G, sk.box);
    // forward: r (natural) -> (q, r) pairs; backward: pairs -> z (natural) and the
    // tiles' dot(z, r) partials
    const W8Planes<T> f{r, sk.oQ, sk.oRec, sk.oEdge, nullptr, nullptr};
    const W8Planes<T> b{sk.oQ, nullptr, sk.oRec + 4 * n, sk.oEdge, z, dot_partial};
    const int tiles = w.ntjb * w.ntkb;
    // (FP32: the deep ring when every tile gets a CU of its own)
    const bool deep = sizeof(T) == 4 && w8_fused_fits<T, sizeof(T) == 4>(tiles);
    sk.fused = deep || w8_fused_fits<T, false>(tiles) ? 1 : 0;
    if (sk.ax == 2) w8_launch<T, 2>(G, w, sk.fused, deep, f, b, S, tag1, tag2, flags, fin, t0, t1, st);
    else w8_launch<T, 0>(G, w, sk.fused, deep, f, b, S, tag1, tag2, flags, fin, t0, t1, st);
    return true;
}

template void sweep_prepare_w8<double>(const Geo&, const uint16_t*, const double*, const SweepSkew<double>&, hipStream_t);
template void sweep_prepare_w8<float>(const Geo&, const uint16_t*, const float*, const SweepSkew<float>&, hipStream_t);
template bool sweep_precond_w8<double>(const Geo&, const SweepSkew<double>&, const double*, double*, const PcgScalars*,
                                       int*, unsigned long long*, hipEvent_t, hipEvent_t, hipStream_t, double*,
                                       FinRes*);
template bool sweep_precond_w8<float>(const Geo&, const SweepSkew<float>&, const float*, float*, const PcgScalars*,
                                      int*, unsigned long long*, hipEvent_t, hipEvent_t, hipStream_t, double*,
                                      FinRes*);
template void w8_dot_tiles<double>(const Geo&, int, const int*, const double*, const double*, double*,
                                   const PcgScalars*, hipStream_t);
template void w8_dot_tiles<float>(const Geo&, int, const int*, const float*, const float*, double*, const PcgScalars*,
                                  hipStream_t);
