// xf_particles.hip -- particle kernels: cell sort, G2P + advection, P2G and
// density gathers, surface classification with prefix-sum compaction.
//
// Particles live as SoA FP64 arrays, kept sorted by (cell, original index)
// after every advection, so that the P2G and density gathers read each cell's
// particles contiguously and sum them in a fixed order (deterministic results,
// tolerance parity with the reference's particle-order sums).
#include "xf_internal.h"

// ---------------------------------------------------------------------------
// Exclusive prefix sum of int32 with wavefront (DPP-shuffle) scans inside each
// block: 256 threads x 4 items per block, then a single-block scan of block
// totals, then the add-back.
#define SCAN_ITEMS 4
#define SCAN_TILE (XF_BLOCK * SCAN_ITEMS)

__device__ inline int32_t wave_incl_scan(int32_t v) {
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        int32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}
// Block exclusive scan of one value per thread; returns the prefix and the total.
__device__ inline int32_t block_excl_scan(int32_t v, int32_t* sh, int32_t* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int32_t inc = wave_incl_scan(v);
    if (lane == 63) sh[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        int32_t s = lane < nw ? sh[lane] : 0;
        int32_t si = wave_incl_scan(s);
        if (lane < nw) sh[lane] = si - s;
        if (lane == nw - 1) sh[16] = si;
    }
    __syncthreads();
    int32_t r = inc - v + sh[wid];
    *total = sh[16];
    __syncthreads();
    return r;
}

__global__ void k_scan_tiles(const int32_t* in, long n, int32_t* tile_sum) {
    __shared__ int32_t sh[17];
    long base = blockIdx.x * (long)SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int32_t s = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) if (base + t < n) s += in[base + t];
    int32_t tot;
    block_excl_scan(s, sh, &tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}
__global__ void k_scan_sums(int32_t* tile_sum, int ntiles) {
    __shared__ int32_t sh[17];
    // one block of 1024: each thread owns a contiguous run
    int per = (ntiles + blockDim.x - 1) / blockDim.x;
    int b0 = threadIdx.x * per;
    int32_t s = 0;
    for (int t = 0; t < per; ++t) if (b0 + t < ntiles) s += tile_sum[b0 + t];
    int32_t tot;
    int32_t pre = block_excl_scan(s, sh, &tot);
    for (int t = 0; t < per; ++t)
        if (b0 + t < ntiles) { int32_t v = tile_sum[b0 + t]; tile_sum[b0 + t] = pre; pre += v; }
}
__global__ void k_scan_apply(const int32_t* in, int32_t* out, long n, const int32_t* tile_sum) {
    __shared__ int32_t sh[17];
    long base = blockIdx.x * (long)SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int32_t v[SCAN_ITEMS];
    int32_t s = 0;
    for (int t = 0; t < SCAN_ITEMS; ++t) { v[t] = base + t < n ? in[base + t] : 0; s += v[t]; }
    int32_t tot;
    int32_t pre = block_excl_scan(s, sh, &tot) + tile_sum[blockIdx.x];
    for (int t = 0; t < SCAN_ITEMS; ++t) {
        if (base + t < n) out[base + t] = pre;
        pre += v[t];
    }
    // out[n] = total, written by the block that owns element n-1
    if (base <= n - 1 && n - 1 < base + SCAN_ITEMS) out[n] = pre;
}

long scan_tmp_size(long n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

// out has n+1 entries; out[n] = sum of all.
void exclusive_scan_i32(const int32_t* in, int32_t* out, long n, int32_t* tmp, int tmp_n, hipStream_t st) {
    int tiles = (int)((n + SCAN_TILE - 1) / SCAN_TILE);
    (void)tmp_n;
    k_scan_tiles<<<tiles, XF_BLOCK, 0, st>>>(in, n, tmp);
    k_scan_sums<<<1, 1024, 0, st>>>(tmp, tiles);
    k_scan_apply<<<tiles, XF_BLOCK, 0, st>>>(in, out, n, tmp);
}

// ---------------------------------------------------------------------------
// Counting sort by cell: key = pos2Index cell (xfluidutil.cpp:3-6) or N when
// outside the grid; within a cell, ascending original index.
// Runs of equal keys over a wavefront's consecutive lanes: particles stored in
// cell order by the previous sort mostly keep their cell, so a run's first lane
// does one atomic for the whole run.  Returns the lane's offset in its run and
// sets the run's first lane and length (for live lanes; the live lanes of a
// wavefront are a prefix).  Every lane of the wavefront must call it.
__device__ inline int key_run(int32_t kk, bool live, int* head, int* len) {
    const int lane = threadIdx.x & 63;
    const int32_t prev = __shfl_up(kk, 1, 64);
    const uint64_t heads = __ballot(live && (lane == 0 || prev != kk));
    const uint64_t alive = __ballot(live);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
    if (!live) {
        *head = lane;
        *len = 0;
        return 0;
    }
    const int hd = 63 - __builtin_clzll(heads & upto);
    const uint64_t stop = (heads | ~alive) & ~upto;
    const int end = stop ? __builtin_ctzll(stop) : 64;
    *head = hd;
    *len = end - hd;
    return lane - hd;
}

__global__ void k_keys(Geo g, long n, const double* px, const double* py, const double* pz, int32_t* key,
                       int32_t* cnt) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const bool live = p < n;
    int32_t kk = -1;
    if (live) {
        int i = pos2idx(px[p], g.dx), j = pos2idx(py[p], g.dx), k = pos2idx(pz[p], g.dx);
        kk = cin(g, i, j, k) ? (int32_t)cidx(g, i, j, k) : (int32_t)g.N;
        key[p] = kk;
    }
    int hd, len;
    const int off = key_run(kk, live, &hd, &len);
    if (live && off == 0) atomicAdd(&cnt[kk], len);
}
// each particle's original index into a free slot of its cell's range
__global__ void k_scatter_ids(long n, const int32_t* __restrict__ key, const int32_t* __restrict__ pid,
                              const int32_t* __restrict__ start, int32_t* cursor, int32_t* ids) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const bool live = p < n;
    const int32_t kk = live ? key[p] : -1;
    int hd, len;
    const int off = key_run(kk, live, &hd, &len);
    int base = 0;
    if (live && off == 0) base = atomicAdd(&cursor[kk], len);
    base = __shfl(base, hd, 64);
    if (live) ids[start[kk] + base + off] = pid[p];
}
// Particle p goes to slot start + (the number of its cell's particles with a
// smaller original index): ascending original index within every cell, as a
// stable sort gives, with no per-cell serial pass (a cell holding m particles
// costs each of them m reads of one contiguous run of ids).
__global__ void k_rank_gather(long n, const int32_t* __restrict__ key, const int32_t* __restrict__ start,
                              const int32_t* __restrict__ ids, const double* px, const double* py, const double* pz,
                              const double* vx, const double* vy, const double* vz, const int32_t* pid, double* tx,
                              double* ty, double* tz, double* tvx, double* tvy, double* tvz, int32_t* tpid) {
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int32_t kk = key[p], a = start[kk], b = start[kk + 1], me = pid[p];
    int32_t r = 0;
    for (int32_t x = a; x < b; ++x) r += ids[x] < me;
    const int32_t s = a + r;
    tx[s] = px[p]; ty[s] = py[p]; tz[s] = pz[p];
    tvx[s] = vx[p]; tvy[s] = vy[p]; tvz[s] = vz[p];
    tpid[s] = me;
}
template <class T>
__global__ void k_zero(long n, T* a) {
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s < n) a[s] = 0;
}

void sort_particles(const Geo& g, Particles& P, CellIndex& ci, hipStream_t st) {
    const long nb = g.N + 1;
    k_zero<int32_t><<<xf_blocks(nb), XF_BLOCK, 0, st>>>(nb, ci.count);
    k_zero<int32_t><<<xf_blocks(nb), XF_BLOCK, 0, st>>>(nb, ci.cursor);
    if (P.n > 0) k_keys<<<xf_blocks(P.n), XF_BLOCK, 0, st>>>(g, P.n, P.px, P.py, P.pz, P.key, ci.count);
    exclusive_scan_i32(ci.count, ci.start, nb, ci.scan_tmp, ci.scan_tmp_n, st);
    if (P.n == 0) return;
    k_scatter_ids<<<xf_blocks(P.n), XF_BLOCK, 0, st>>>(P.n, P.key, P.pid, ci.start, ci.cursor, P.perm);
    k_rank_gather<<<xf_blocks(P.n), XF_BLOCK, 0, st>>>(P.n, P.key, ci.start, P.perm, P.px, P.py, P.pz, P.vx, P.vy,
                                                       P.vz, P.pid, P.tx, P.ty, P.tz, P.tvx, P.tvy, P.tvz, P.tpid);
    std::swap(P.px, P.tx); std::swap(P.py, P.ty); std::swap(P.pz, P.tz);
    std::swap(P.vx, P.tvx); std::swap(P.vy, P.tvy); std::swap(P.vz, P.tvz);
    std::swap(P.pid, P.tpid);
}


// _advectParticles (FluidSimulator.cpp:652-723): FLIP/PIC blend, RK4 (or RK2
// midpoint) advection, collision resolution.  Particles whose original index
// is >= limit_id are left untouched (the reference's 6-chunk tail).
__global__ void k_g2p_advect(Geo g, long n, double* px, double* py, double* pz, double* vx, double* vy, double* vz,
                             const int32_t* pid, const uint8_t* mat, Faces Fn, Faces Fo, double dt, double flip,
                             int integrator, long limit_id) {
    long p = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (pid[p] >= limit_id) return;
    double pos[3] = {px[p], py[p], pz[p]};
    double vn[3], vo[3];
    sample(g, Fn, pos, vn);
    sample(g, Fo, pos, vo);
    double vp[3] = {vx[p], vy[p], vz[p]};
    for (int a = 0; a < 3; ++a) {
        double vpic = vn[a];
        double vflip = vp[a] + vn[a] - vo[a];
        vp[a] = vpic * (1.0 - flip) + vflip * flip;
    }
    vx[p] = vp[0]; vy[p] = vp[1]; vz[p] = vp[2];

    double np_[3], q[3], k1[3], k2[3];
    sample(g, Fn, pos, k1);
    for (int a = 0; a < 3; ++a) q[a] = pos[a] + k1[a] * (0.5 * dt);
    sample(g, Fn, q, k2);
    if (integrator == 1) {
        for (int a = 0; a < 3; ++a) np_[a] = pos[a] + k2[a] * dt;
    } else {
        double k3[3], k4[3];
        for (int a = 0; a < 3; ++a) q[a] = pos[a] + k2[a] * (0.5 * dt);
        sample(g, Fn, q, k3);
        for (int a = 0; a < 3; ++a) q[a] = pos[a] + k3[a] * dt;
        sample(g, Fn, q, k4);
        for (int a = 0; a < 3; ++a) np_[a] = pos[a] + (k1[a] + k2[a] * 2.0 + k3[a] * 2.0 + k4[a]) * (dt / 6.0);
    }
    double out[3];
    Collision c;
    if (resolve_move(g, mat, pos, np_, out, &c) != MOVE_KEPT) {
        px[p] = out[0]; py[p] = out[1]; pz[p] = out[2];
    }
}

// Particles as n x 3 interleaved arrays in id order (the order they were set
// in): gather from / scatter to the sorted SoA arrays by their ids.
__global__ void k_to_n3(long n, const double* __restrict__ px, const double* __restrict__ py,
                        const double* __restrict__ pz, const double* __restrict__ vx, const double* __restrict__ vy,
                        const double* __restrict__ vz, const int32_t* __restrict__ pid, double* pos, double* vel) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const long q = 3 * (long)pid[p];
    if (pos) { pos[q] = px[p]; pos[q + 1] = py[p]; pos[q + 2] = pz[p]; }
    if (vel) { vel[q] = vx[p]; vel[q + 1] = vy[p]; vel[q + 2] = vz[p]; }
}
__global__ void k_from_n3(long n, const double* __restrict__ pos, const double* __restrict__ vel, double* px,
                          double* py, double* pz, double* vx, double* vy, double* vz, int32_t* pid) {
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= n) return;
    px[p] = pos[3 * p]; py[p] = pos[3 * p + 1]; pz[p] = pos[3 * p + 2];
    vx[p] = vel[3 * p]; vy[p] = vel[3 * p + 1]; vz[p] = vel[3 * p + 2];
    pid[p] = (int32_t)p;
}
void particles_to_n3(const Particles& P, double* pos, double* vel, hipStream_t st) {
    if (P.n > 0) k_to_n3<<<xf_blocks(P.n), XF_BLOCK, 0, st>>>(P.n, P.px, P.py, P.pz, P.vx, P.vy, P.vz, P.pid, pos, vel);
}
void particles_from_n3(Particles& P, const double* pos, const double* vel, hipStream_t st) {
    if (P.n > 0)
        k_from_n3<<<xf_blocks(P.n), XF_BLOCK, 0, st>>>(P.n, pos, vel, P.px, P.py, P.pz, P.vx, P.vy, P.vz, P.pid);
}

void launch_g2p_advect(const Geo& g, Particles& P, const uint8_t* mat, const double* u, const double* v,
                       const double* w, const double* us, const double* vs, const double* ws, double dt, double flip,
                       int integrator, int64_t limit_id, hipStream_t st) {
    if (P.n == 0) return;
    Faces Fn{u, v, w}, Fo{us, vs, ws};
    k_g2p_advect<<<xf_blocks(P.n), XF_BLOCK, 0, st>>>(g, P.n, P.px, P.py, P.pz, P.vx, P.vy, P.vz, P.pid, mat, Fn, Fo,
                                                       dt, flip, integrator, limit_id);
}

// ---------------------------------------------------------------------------
// CubicFluidSource::update (fluidsource.cpp:75-92), applied at the end of
// _advectParticles (FluidSimulator.cpp:725-726).  Thread t of one source's
// launch writes particle n0 + t: cell cells[t / per_cell], position = cell
// centre (idx2Pos, xfluidutil.cpp:18-21) + randomDouble(-0.4999 dx, 0.4999 dx)
// per axis (xfluidutil.cpp:67-70), velocity = the source's, id = n0 + t.  The
// draws come from a 31-bit SplitMix64 stream in place of rand() (RAND_MAX =
// 2^31 - 1): draw d is mix(seed + (d + 1) * 0x9E3779B97F4A7C15), so thread t
// takes draws 3 (m0 + t) + 0, 1, 2 (x, y, z) without walking the stream.
__device__ inline uint32_t source_draw(uint64_t seed, uint64_t d) {
    uint64_t z = seed + (d + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (uint32_t)((z ^ (z >> 31)) >> 33);
}

__global__ void k_source_emit(Geo g, const int32_t* __restrict__ cells, long total, int per_cell, double svx,
                              double svy, double svz, uint64_t seed, uint64_t m0, long n0, double* px, double* py,
                              double* pz, double* vx, double* vy, double* vz, int32_t* pid) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const long c = cells[t / per_cell];
    int i, j, k;
    unflat3(c, g.ni, g.nj, &i, &j, &k);
    const double lo = -0.4999 * g.dx, hi = 0.4999 * g.dx;
    const double scale = 2147483647.0 / (hi - lo);
    const uint64_t d = 3 * (m0 + (uint64_t)t);
    const double jx = lo + (double)source_draw(seed, d) / scale;
    const double jy = lo + (double)source_draw(seed, d + 1) / scale;
    const double jz = lo + (double)source_draw(seed, d + 2) / scale;
    const long q = n0 + t;
    px[q] = (i + 0.5) * g.dx + jx;
    py[q] = (j + 0.5) * g.dx + jy;
    pz[q] = (k + 0.5) * g.dx + jz;
    vx[q] = svx; vy[q] = svy; vz[q] = svz;
    pid[q] = (int32_t)q;
}

__global__ void k_source_mark(uint8_t* mat, const int32_t* __restrict__ cells, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) mat[cells[t]] = XF_FLUID;
}

void launch_source(const Geo& g, Particles& P, uint8_t* mat, const int32_t* cells, int ncells, int per_cell,
                   const double v[3], uint64_t seed, uint64_t m0, int64_t n0, hipStream_t st) {
    if (ncells <= 0) return;
    k_source_mark<<<(ncells + XF_BLOCK - 1) / XF_BLOCK, XF_BLOCK, 0, st>>>(mat, cells, ncells);
    const long total = (long)ncells * per_cell;
    if (total <= 0) return;
    k_source_emit<<<(unsigned)((total + XF_BLOCK - 1) / XF_BLOCK), XF_BLOCK, 0, st>>>(
        g, cells, total, per_cell, v[0], v[1], v[2], seed, m0, n0, P.px, P.py, P.pz, P.vx, P.vy, P.vz, P.pid);
}

// ---------------------------------------------------------------------------
// P2G, _updateVelocityByParticles (FluidSimulator.cpp:909-1085), as a gather
// over bricks of 8x8x4 faces: face f = sum_p W(|q_p - c_f|^2) v_p / sum_p W,
// q_p = p + dx/2 e_axis, c_f = idx2Pos(f); faces with sum W <= 1e-7 get 0.  A
// workgroup first reads the cell starts of all its candidate slabs (one round
// of loads), then walks the k-slabs that hold particles: it copies the slab's
// particles (one contiguous run per cell row of the sorted arrays) into LDS
// with several loads in flight per thread, then every face adds the slab's
// rows inside its support, reading particles from LDS, cells in k, j, i order
// and particles by index.  Each particle is fetched once per brick instead of
// once per face, and bricks in air cost one round of loads.  A slab with more
// particles than fit is staged in pieces of P2G_CAP particles, which keeps the
// same order.  (Reading every face's particles straight from global memory
// instead took twice as long on the 128^3 dam break.)  Only the
// summation order differs from the reference's, so the result equals it to
// rounding (the weights may use fused multiply-adds).
#ifndef P2G_BJ
#define P2G_BJ 8
#endif
#ifndef P2G_BK
#define P2G_BK 4
#endif
#ifndef P2G_CAP
#define P2G_CAP 1024  // 34 KB of staged particles: 4 workgroups (16 waves) per CU
#endif
#define P2G_BI 8
#define P2G_MAXC 12  // cells along i in a slab row (8 faces + support)
// (KERN: 0 Wyvill, 1 trilinear; AX: the face axis; one kernel each, so the pair
// loop carries only its own weight's registers and shifts only the axis component)
template <int KERN, int AX>
__global__ __launch_bounds__(P2G_BI * P2G_BJ * P2G_BK) void k_p2g_brick(
    Geo g, int nbi, int nbj, const double* __restrict__ px, const double* __restrict__ py,
    const double* __restrict__ pz, const double* __restrict__ vc, const int32_t* __restrict__ start,
    double* __restrict__ out) {
#pragma clang fp contract(fast)  // a tolerance kernel: its sums run in another order than the reference's anyway
    typedef double d2 __attribute__((ext_vector_type(2)));
    // slab particle q: {x, y} at 2q + q/8, {z, v} after it.  A cell's particles
    // follow one another and most cells hold 8, so lanes one face apart read 8
    // particles (256 bytes) apart: without the 16-byte step after every 8
    // particles they would all read the same LDS banks.
    __shared__ d2 lp[2 * P2G_CAP + P2G_CAP / 8];
    __shared__ int raw[P2G_BK + 3][P2G_BJ + 3][P2G_MAXC + 1];  // start[] of the slabs' cells, per row
    __shared__ int roff[P2G_BK + 3][P2G_BJ + 4];               // LDS offset of each slab row, then the slab's total
    constexpr int axis = AX;
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1), sk = g.nk + (axis == 2);
    const long b = xcd_block(blockIdx.x, gridDim.x);
    const int bi = (int)(b % nbi), bj = (int)((b / nbi) % nbj), bk = (int)(b / ((long)nbi * nbj));
    const int F[3] = {bi * P2G_BI, bj * P2G_BJ, bk * P2G_BK};
    const int tid = threadIdx.x;
    const int fi = F[0] + tid % P2G_BI, fj = F[1] + (tid / P2G_BI) % P2G_BJ, fk = F[2] + tid / (P2G_BI * P2G_BJ);
    const bool valid = fi < si && fj < sj && fk < sk;
    constexpr bool tri = KERN == 1;
    const double dx = g.dx;
    const double r = tri ? dx : 1.5 * dx;
    const Wyvill W = make_wyvill(r);
    const int n[3] = {g.ni, g.nj, g.nk};
    // this face's support cells and the brick's candidate cells, per direction
    int lo[3], hi[3], clo[3], chi[3];
    const int fc[3] = {fi, fj, fk}, bsz[3] = {P2G_BI, P2G_BJ, P2G_BK};
    for (int a = 0; a < 3; ++a) {
        const int dl = a == axis ? (tri ? 1 : 2) : 1, dh = a == axis ? (tri ? 0 : 1) : 1;
        lo[a] = fc[a] - dl; hi[a] = fc[a] + dh;
        clo[a] = F[a] - dl; chi[a] = F[a] + bsz[a] - 1 + dh;
        if (lo[a] < 0) lo[a] = 0;
        if (hi[a] > n[a] - 1) hi[a] = n[a] - 1;
        if (clo[a] < 0) clo[a] = 0;
        if (chi[a] > n[a] - 1) chi[a] = n[a] - 1;
    }
    const double cx = idx2pos(fi, dx), cy = idx2pos(fj, dx), cz = idx2pos(fk, dx);
    const double hdx = 0.5 * dx;
    double wsum = 0.0, vsum = 0.0;
    auto pair = [&](double qx, double qy, double qz, double v) __attribute__((always_inline)) {
        // q = p + dx/2 e_axis (the other components' + 0 changes no value)
        const double ex = (AX == 0 ? qx + hdx : qx) - cx, ey = (AX == 1 ? qy + hdx : qy) - cy,
                     ez = (AX == 2 ? qz + hdx : qz) - cz;
        double cw;
        if constexpr (tri) {
            const double ax = fabs(ex) / dx, ay = fabs(ey) / dx, az = fabs(ez) / dx;
            cw = (ax < 1.0 && ay < 1.0 && az < 1.0) ? (1.0 - ax) * (1.0 - ay) * (1.0 - az) : 0.0;
        } else {
            const double s2 = ex * ex + ey * ey + ez * ez;
            cw = wyvill(W, s2);
        }
        // (a pair outside the support has weight +0, which leaves both sums' values as
        // they are: no branch)
        wsum += cw;
        vsum += v * cw;
    };
    const int nrow = chi[1] - clo[1] + 1, ncell = chi[0] - clo[0] + 1, nsl = chi[2] - clo[2] + 1;
    // the cell starts of every candidate slab row
    const int per = nrow * (ncell + 1);
    for (int q = tid; q < nsl * per; q += blockDim.x) {
        const int ks = q / per, rem = q - ks * per, jr = rem / (ncell + 1), ic = rem - jr * (ncell + 1);
        raw[ks][jr][ic] = start[(long)g.ni * (clo[1] + jr) + g.nn * (clo[2] + ks) + clo[0] + ic];
    }
    __syncthreads();
    if (tid < nsl) {
        int acc = 0;
        for (int jr = 0; jr < nrow; ++jr) {
            roff[tid][jr] = acc;
            acc += raw[tid][jr][ncell] - raw[tid][jr][0];
        }
        roff[tid][nrow] = acc;
    }
    __syncthreads();
    constexpr int NB = P2G_BI * P2G_BJ * P2G_BK, UNR = 4;
    static_assert(UNR * NB >= P2G_CAP, "one staging round per chunk");
    for (int ks = 0; ks < nsl; ++ks) {
        const int k = clo[2] + ks;
        const int m = roff[ks][nrow];
        // (uniform) the slab's particles in chunks of at most P2G_CAP, in slab
        // order: a face adds its rows' particles chunk by chunk, which is the
        // same order (rows, then particle index) as in one piece
        for (int c0 = 0; c0 < m; c0 += P2G_CAP) {
            const int c1 = m - c0 < P2G_CAP ? m : c0 + P2G_CAP;
            // slab particle q: row jr with roff[jr] <= q < roff[jr + 1]
            {
                double vx[UNR], vy[UNR], vz[UNR], vv[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    const int q = c0 + tid + u * NB;
                    if (q < c1) {
                        int jr = 0;
                        while (jr + 1 < nrow && roff[ks][jr + 1] <= q) ++jr;
                        const int a = raw[ks][jr][0] + (q - roff[ks][jr]);
                        vx[u] = px[a]; vy[u] = py[a]; vz[u] = pz[a]; vv[u] = vc[a];
                    }
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    const int q = c0 + tid + u * NB;
                    if (q < c1) {
                        const int e = 2 * (q - c0) + ((q - c0) >> 3);
                        lp[e] = d2{vx[u], vy[u]};
                        lp[e + 1] = d2{vz[u], vv[u]};
                    }
                }
            }
            __syncthreads();
            if (valid && k >= lo[2] && k <= hi[2]) {
                for (int j = lo[1]; j <= hi[1]; ++j) {
                    const int jr = j - clo[1];
                    const int base = roff[ks][jr] - raw[ks][jr][0];
                    int pa = base + raw[ks][jr][lo[0] - clo[0]], pb = base + raw[ks][jr][hi[0] + 1 - clo[0]];
                    pa = (pa > c0 ? pa : c0) - c0;
                    pb = (pb < c1 ? pb : c1) - c0;
                    // four particles' LDS reads in flight at once, then added in index order
                    int p = pa;
                    for (; p + 4 <= pb; p += 4) {
                        d2 a[4], c[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int e = 2 * (p + u) + ((p + u) >> 3);
                            a[u] = lp[e];
                            c[u] = lp[e + 1];
                        }
#pragma unroll
                        for (int u = 0; u < 4; ++u) pair(a[u].x, a[u].y, c[u].x, c[u].y);
                    }
                    for (; p < pb; ++p) {
                        const int e = 2 * p + (p >> 3);
                        const d2 a = lp[e], c = lp[e + 1];
                        pair(a.x, a.y, c.x, c.y);
                    }
                }
            }
            __syncthreads();
        }
    }
    if (valid) out[fi + (long)si * (fj + (long)sj * fk)] = wsum > 0.0000001 ? vsum / wsum : 0.0;
}

void launch_p2g(const Geo& g, const Particles& P, const CellIndex& ci, double* f, int axis, int kernel,
                hipStream_t st) {
    const int si = g.ni + (axis == 0), sj = g.nj + (axis == 1), sk = g.nk + (axis == 2);
    const double* vc = axis == 0 ? P.vx : axis == 1 ? P.vy : P.vz;
    const int nbi = (si + P2G_BI - 1) / P2G_BI, nbj = (sj + P2G_BJ - 1) / P2G_BJ, nbk = (sk + P2G_BK - 1) / P2G_BK;
    void (*const ks[2][3])(Geo, int, int, const double*, const double*, const double*, const double*, const int32_t*,
                           double*) = {{k_p2g_brick<0, 0>, k_p2g_brick<0, 1>, k_p2g_brick<0, 2>},
                                       {k_p2g_brick<1, 0>, k_p2g_brick<1, 1>, k_p2g_brick<1, 2>}};
    ks[kernel == 1][axis]<<<nbi * nbj * nbk, P2G_BI * P2G_BJ * P2G_BK, 0, st>>>(g, nbi, nbj, P.px, P.py, P.pz, vc,
                                                                              ci.start, f);
}

// ---------------------------------------------------------------------------
// Particle-blob density, DirectedDistanceField::updateFeild
// (directeddistancefield.cpp:122-169, 300-328): phi(v) = sum_p W(|x_v - p|^2)
// over the (d*n+1)^3 sub-vertex lattice, x_v = (v + 1/2) dx/d.  One thread
// gathers the d^3 vertices owned by one coarse cell, over the particles of the
// cells within R of them (rows of the cell-sorted arrays, in k, j, particle
// index order); the per-axis squared offsets are computed once per particle.
// A workgroup is one wavefront of 8x8 coarse cells of one k layer, so
// neighbouring lanes read overlapping runs of particles through the caches.
// (Staging each k-slab's particles in LDS instead, for bricks of 8x8x4 cells
// with a barrier per slab, left the wavefronts that do not need a slab idle at
// its barriers: 740 against 464 us on the 128^3 dam break.)  Only the summation
// order and the kernel polynomial's form (Horner, with fused multiply-adds)
// differ from the reference, so the result equals it to rounding.
#define DB_BI 8
#define DB_BJ 8
#define DB_BK 1
template <int D>
__global__ __launch_bounds__(DB_BI * DB_BJ * DB_BK) void k_density(Geo g, int nbi, int nbj,
                                                                   const double* __restrict__ px,
                                                                   const double* __restrict__ py,
                                                                   const double* __restrict__ pz,
                                                                   const int32_t* __restrict__ start,
                                                                   double* __restrict__ phi, double R) {
#pragma clang fp contract(fast)  // a tolerance kernel: its sums run in another order than the reference's anyway
    const long b = xcd_block(blockIdx.x, gridDim.x);
    const int bi = (int)(b % nbi), bj = (int)((b / nbi) % nbj), bk = (int)(b / ((long)nbi * nbj));
    const int tid = threadIdx.x;
    const int ci = bi * DB_BI + tid % DB_BI, cj = bj * DB_BJ + (tid / DB_BI) % DB_BJ,
              ck = bk * DB_BK + tid / (DB_BI * DB_BJ);
    if (ci > g.ni || cj > g.nj || ck > g.nk) return;
    const int vi = D * g.ni + 1, vj = D * g.nj + 1;
    const int ni_v = (ci == g.ni) ? 1 : D, nj_v = (cj == g.nj) ? 1 : D, nk_v = (ck == g.nk) ? 1 : D;
    const double dx = g.dx, subdx = dx / (double)D;
    const Wyvill W = make_wyvill(R);
    double xs[D], ys[D], zs[D];
    for (int a = 0; a < D; ++a) {
        xs[a] = idx2pos(D * ci + a, subdx);
        ys[a] = idx2pos(D * cj + a, subdx);
        zs[a] = idx2pos(D * ck + a, subdx);
    }
    // candidate cells: those overlapping [x_lo - R, x_hi + R] (tiny margin for rounding at cell edges)
    auto range = [&](double lo_pos, double hi_pos, int n, int* lo, int* hi) {
        *lo = (int)floor((lo_pos - R) / dx - 1e-7);
        *hi = (int)floor((hi_pos + R) / dx + 1e-7);
        if (*lo < 0) *lo = 0;
        if (*hi > n - 1) *hi = n - 1;
    };
    int li, hi_, lj, hj, lk, hk;
    range(xs[0], xs[ni_v - 1], g.ni, &li, &hi_);
    range(ys[0], ys[nj_v - 1], g.nj, &lj, &hj);
    range(zs[0], zs[nk_v - 1], g.nk, &lk, &hk);
    double acc[D][D][D];
    for (int a = 0; a < D; ++a) for (int c = 0; c < D; ++c) for (int e = 0; e < D; ++e) acc[a][c][e] = 0.0;
    auto one = [&](double x, double y, double z) __attribute__((always_inline)) {
        double ex2[D], ey2[D], ez2[D];
        for (int a = 0; a < D; ++a) {
            double e = xs[a] - x; ex2[a] = e * e;
            e = ys[a] - y; ey2[a] = e * e;
            e = zs[a] - z; ez2[a] = e * e;
        }
        // the kernel polynomial in Horner form, y + z shared by the row of
        // vertices: nothing divides by these sums (unlike P2G's weight sums), so
        // the form's rounding stays at ~1e-16 of the kernel's peak
        for (int c = 0; c < D; ++c)
            for (int bb = 0; bb < D; ++bb) {
                const double yz = ey2[bb] + ez2[c];
                for (int a = 0; a < D; ++a) {
                    const double s = ex2[a] + yz;
                    const double w = fma(fma(fma(-W.c0, s, W.c1), s, -W.c2), s, 1.0);
                    acc[c][bb][a] += s < W.r2 ? w : 0.0;
                }
            }
    };
    for (int k = lk; k <= hk; ++k)
        for (int j = lj; j <= hj; ++j) {
            const long row = (long)g.ni * j + g.nn * k;
            const int a0 = start[row + li], b0 = start[row + hi_ + 1];
            // four particles' positions requested together, then added in index order
            int p = a0;
            for (; p + 4 <= b0; p += 4) {
                double X[4], Y[4], Z[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) { X[u] = px[p + u]; Y[u] = py[p + u]; Z[u] = pz[p + u]; }
#pragma unroll
                for (int u = 0; u < 4; ++u) one(X[u], Y[u], Z[u]);
            }
            for (; p < b0; ++p) one(px[p], py[p], pz[p]);
        }
    for (int c = 0; c < nk_v; ++c)
        for (int bb = 0; bb < nj_v; ++bb)
            for (int a = 0; a < ni_v; ++a) {
                const long v = (D * ci + a) + (long)vi * ((D * cj + bb) + (long)vj * (D * ck + c));
                phi[v] = acc[c][bb][a];
            }
}

void launch_density(const Geo& g, const Particles& P, const CellIndex& ci, double* phi, int d, double radius,
                    hipStream_t st) {
    const int nbi = (g.ni + DB_BI) / DB_BI, nbj = (g.nj + DB_BJ) / DB_BJ, nbk = (g.nk + DB_BK) / DB_BK;
    const int nb = nbi * nbj * nbk, nt = DB_BI * DB_BJ * DB_BK;
    if (d == 1) k_density<1><<<nb, nt, 0, st>>>(g, nbi, nbj, P.px, P.py, P.pz, ci.start, phi, radius);
    else if (d == 2) k_density<2><<<nb, nt, 0, st>>>(g, nbi, nbj, P.px, P.py, P.pz, ci.start, phi, radius);
    else if (d == 3) k_density<3><<<nb, nt, 0, st>>>(g, nbi, nbj, P.px, P.py, P.pz, ci.start, phi, radius);
    else k_density<4><<<nb, nt, 0, st>>>(g, nbi, nbj, P.px, P.py, P.pz, ci.start, phi, radius);
}

// ---------------------------------------------------------------------------
// Surface cells (directeddistancefield.cpp:171-244, 330-352) and cube case
// (fluidmesher.cpp:397-417).  Coarse cells are ranked in the reference's loop
// order (i slowest, k fastest); a coarse cell is a surface cell if its
// (d+1)^3 vertices straddle iso (>= in / < out); its sub-cells, in (a,b,c)
// order with c fastest, are kept if their 8 corners straddle iso.  Pass 1
// counts per coarse cell, a prefix sum ranks them, pass 2 writes.
__device__ inline long coarse_of_rank(const Geo& g, long t, int* i, int* j, int* k) {
    *k = (int)(t % g.nk);
    *j = (int)((t / g.nk) % g.nj);
    *i = (int)(t / ((long)g.nk * g.nj));
    return t;
}

template <bool WRITE>
__global__ void k_surface(Geo g, int d, const double* __restrict__ phi, double iso, int32_t* counts,
                          const int32_t* offs, int32_t* cells, uint8_t* cube, long cap) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= g.N) return;
    int i, j, k;
    coarse_of_rank(g, t, &i, &j, &k);
    const long vi = d * g.ni + 1, vj = d * g.nj + 1;
    auto PHI = [&](int a, int b, int c) { return phi[a + vi * (b + vj * c)]; };
    const int s0 = i * d, s1 = j * d, s2 = k * d;
    bool in = false, out = false;
    for (int a = s0; a <= s0 + d; ++a)
        for (int b = s1; b <= s1 + d; ++b)
            for (int c = s2; c <= s2 + d; ++c) {
                if (PHI(a, b, c) >= iso) in = true; else out = true;
            }
    int cnt = 0;
    long pos = WRITE ? offs[t] : 0;
    if (in && out) {
        const long sni = (long)d * g.ni, snj = (long)d * g.nj;
        for (int a = s0; a < s0 + d; ++a)
            for (int b = s1; b < s1 + d; ++b)
                for (int c = s2; c < s2 + d; ++c) {
                    // vs[] corner order of fluidmesher.cpp:397-406
                    double v0 = PHI(a, b, c), v1 = PHI(a + 1, b, c), v2 = PHI(a + 1, b, c + 1), v3 = PHI(a, b, c + 1);
                    double v4 = PHI(a, b + 1, c), v5 = PHI(a + 1, b + 1, c), v6 = PHI(a + 1, b + 1, c + 1),
                           v7 = PHI(a, b + 1, c + 1);
                    bool ii = false, oo = false;
                    double vv[8] = {v0, v1, v2, v3, v4, v5, v6, v7};
                    for (int q = 0; q < 8; ++q) { if (vv[q] >= iso) ii = true; else oo = true; }
                    if (!(ii && oo)) continue;
                    if (WRITE && pos + cnt < cap) {
                        int ci = 0;
                        for (int q = 0; q < 8; ++q) if (vv[q] > iso) ci |= 1 << q;
                        cells[pos + cnt] = (int32_t)(a + sni * (b + snj * c));
                        cube[pos + cnt] = (uint8_t)ci;
                    }
                    ++cnt;
                }
    }
    if (!WRITE) counts[t] = cnt;
}

int64_t surface_cells(const Geo& g, const double* phi, int d, double iso, int32_t* cells, uint8_t* cube,
                      int64_t cap, int32_t* count_tmp, int32_t* start_tmp, int32_t* scan_tmp, int scan_tmp_n,
                      hipStream_t st) {
    k_surface<false><<<xf_blocks(g.N), XF_BLOCK, 0, st>>>(g, d, phi, iso, count_tmp, nullptr, nullptr, nullptr, 0);
    exclusive_scan_i32(count_tmp, start_tmp, g.N, scan_tmp, scan_tmp_n, st);
    int32_t total = 0;
    hipMemcpyAsync(&total, start_tmp + g.N, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    if (cells && cap > 0)
        k_surface<true><<<xf_blocks(g.N), XF_BLOCK, 0, st>>>(g, d, phi, iso, nullptr, start_tmp, cells, cube, cap);
    return total;
}

// ---------------------------------------------------------------------------
// max |v|^2 over particles (FluidSimulator.cpp:387-400): v*v = (vx*vx + vy*vy) + vz*vz
__global__ void k_maxv2(long n, const double* vx, const double* vy, const double* vz, double* partial) {
    __shared__ double sh[16];
    double m = 0.0;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        double c = vx[p] * vx[p] + vy[p] * vy[p] + vz[p] * vz[p];
        m = c > m ? c : m;
    }
    m = block_reduce<double, true>(m, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}
void launch_maxv2(const Particles& P, double* partial, int nblocks, hipStream_t st) {
    k_maxv2<<<nblocks, XF_BLOCK, 0, st>>>(P.n, P.vx, P.vy, P.vz, partial);
}
