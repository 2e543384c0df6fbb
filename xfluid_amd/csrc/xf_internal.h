// xf_internal.h -- host-side launchers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xf_device.h"

#define XF_BLOCK 256

inline unsigned xf_blocks(long n, int per = XF_BLOCK) { return (unsigned)((n + per - 1) / per); }

// Scalars of one PCG solve, resident in device memory (no host round trip per
// iteration).  `done` stops all later kernels of the solve.
struct PcgScalars {
    double rho, sz, alpha, beta, res, tol, rho_new, res0;
    double rho_next;  // rho of the next iteration, moved into rho by the next spmv
    int done, iters, status, it;
};
// The residual test of PCG iteration `it` (pcg_solver.h:266-270): max of the n
// partials of |r|, then S->res, it, done and iters, and S copied to `mirror`
// when given.  Handed to the preconditioner that follows it, which runs it
// (inside the fused sweep kernel, or as k_finish before its sweeps) and sets `ran`.
struct FinRes {
    const double* partial;
    int n, it;
    PcgScalars* S;
    PcgScalars* mirror;
    bool ran;
};
void launch_finish_res(const FinRes& f, hipStream_t st);

// ---------------- grid stage kernels (xf_grid.hip) ----------------
void launch_material(const Geo& g, uint8_t* mat, const int32_t* cell_count, hipStream_t st);
void launch_forces(const Geo& g, const uint8_t* mat, double* f, int axis, double dv, hipStream_t st);
// op: OP_VARIATIONAL or OP_NORMAL
void launch_stencil_rhs(const Geo& g, const uint8_t* mat, const double* u, const double* v, const double* w,
                        uint16_t* code, double* b, hipStream_t st, int op = OP_VARIATIONAL);
// the three face arrays in one launch each (per face as the reference's
// per-axis loops)
void launch_apply_pressure(const Geo& g, const uint8_t* mat, const double* p, double* fu, double* fv, double* fw,
                           hipStream_t st, int op);
// marker: three arrays of mstride bytes (one per face array)
void launch_extrapolate(const Geo& g, const uint8_t* mat, double* fu, double* fv, double* fw, int8_t* marker,
                        long mstride, int layers, hipStream_t st);
// VJP of launch_extrapolate for gradient gf on one face array (in place)
void launch_extrapolate_adjoint(const Geo& g, const uint8_t* mat, double* gf, int8_t* marker, int axis, int layers,
                                hipStream_t st);

// ---------------- pressure solve kernels (xf_pcg.hip) ----------------
// Buffers of the pipelined 8x8-tile sweeps (xf_sweep_wave.hip), oct layout:
// forward and backward records, the forward output ((q, r) pairs: 2 values per
// element), tagged edge words.
template <class T>
struct SweepSkew {
    T* oRec;
    T* oQ;
    uint32_t* oEdge;
    mutable uint32_t w8_launch;  // launches on oEdge (their tags)
    mutable int fused;           // last application: 1 one fused forward + backward launch, 0 two launches
    int ax;                      // axis of the tiles' lines: 0 along i, 2 along k (w8_choose_axis)
    // bounding box of the active rows (i0, i1, j0, j1, k0, k1; host copy), see pcg_active_box
    int box[6];
};
// Optional device timing of the sweep launches: events on the forward +
// backward launches of every SWEEP_TIMER_STRIDE-th preconditioner application
// of a solve (application k of the solve, k % stride == 0; timing every
// application would slow the solve it measures).  The stride is prime to the
// 4-iteration chunks, so the samples cover every position in a chunk alike,
// the chunk's last launch (which also writes the host status slot) included.
#define SWEEP_TIMER_STRIDE 5
struct SweepTimer {
    hipEvent_t* ev;
    int cap, used;
    bool on;
    int k;  // applications so far in the current solve
};
struct PcgWork {
    // FP64 vectors
    double *x, *r, *s, *z, *q, *b, *inv;
    double* s2;  // second search-direction buffer (s alternates between s and s2)
    // FP32 copies for the mixed-precision inner solve
    float *x32, *r32, *s32, *z32, *q32, *inv32;
    float* s2_32;
    double* partial;     // per-block partial sums / maxima: two halves of npartial
    int npartial;        // (at least the reduction blocks and the tiles of any box)
    PcgScalars* S;       // device scalars
    PcgScalars* S_host;  // pinned mirror: [0] inner solve, [1] outer, [2] inner (odd chunks)
    hipEvent_t chunk_ev[2];  // ends of the PCG's iteration chunks in flight
    int* sweep_flags;    // error word + progress counters of the pipelined sweep
    int* err_host;       // pinned copy of the error word, read after a solve
    int nflags;
    SweepSkew<double> sk64;  // pipelined sweep buffers (FP64 solve)
    SweepSkew<float> sk32;   // (FP32 inner solve)
    SweepTimer timer;
    unsigned long long epoch;  // sweep launch counter (tags progress counters)
    int* box_dev;   // 6 ints: bounding box of the active rows
    int* box_host;  // pinned mirror
};
struct PcgResult {
    int status, iterations, outer;
    double residual, rhs_norm;
    long box_cells;  // cells of the active rows' bounding box
    // err_known: sweep_err holds the pipelined sweeps' error word as the solve
    // ended (the mixed solve reads it with its last residual; otherwise the
    // caller reads it)
    int sweep_err;
    bool err_known;
};
// Assemble the stencil codes and right-hand side, factor, and solve A x = b.
PcgResult pcg_solve(const Geo& g, const uint16_t* code, PcgWork& w, double tol_factor, int max_iter, double tau,
                    double sigma, int precision, double inner_tol, int max_outer, int sweep_kernel, hipStream_t st);
// Individual pieces, for parity checks.
void pcg_mic0(const Geo& g, const uint16_t* code, double* inv, double tau, double sigma, const int* box,
              hipStream_t st);
// Bounding box of the active rows (rows with a diagonal or an off-diagonal
// entry); false when there are none.  Rows outside it are empty: their
// pivots, residuals and preconditioned values are zero.
bool pcg_active_box(const Geo& g, const uint16_t* code, PcgWork& w, int box[6], hipStream_t st);
void pcg_precond(const Geo& g, const uint16_t* code, const double* inv, const double* r, double* q, double* z,
                 int sweep_kernel, PcgWork& w, hipStream_t st);
void pcg_spmv(const Geo& g, const uint16_t* code, const double* x, double* y, hipStream_t st);
// ints of the pipelined sweep's flags buffer (flags[0]: error word)
int sweep_flag_words(const Geo& g);
// MIC(0) pivots through the 8x8-tile pipeline over sk.box (false: not run)
bool mic0_w8(const Geo& g, const uint16_t* code, double* inv, double tau, double sigma, const SweepSkew<double>& sk,
             int* flags, unsigned long long* epoch, hipStream_t st);
long sweep_w8_size(const Geo& g);
long sweep_w8_edge_words(const Geo& g);
template <class T>
void sweep_prepare_w8(const Geo& g, const uint16_t* code, const T* inv, const SweepSkew<T>& sk, hipStream_t st);
// z = M^-1 r over sk.box, z written at the box's tiles' cells (natural layout;
// other cells untouched).  dot_partial: also the tiles' partials of dot(z, r),
// w8_box_tiles of them, in w8_dot_tiles' order.
template <class T>
bool sweep_precond_w8(const Geo& g, const SweepSkew<T>& sk, const T* r, T* z, const PcgScalars* S, int* flags,
                      unsigned long long* epoch, hipEvent_t t0, hipEvent_t t1, hipStream_t st,
                      double* dot_partial = nullptr, FinRes* fin = nullptr);
// tiles of the 8x8-tile schedule with lines along axis ax over a box (i0, i1, j0, j1, k0, k1)
int w8_box_tiles(const Geo& g, int ax, const int* box);
// the axis (0: i, 2: k) whose lines give the shortest pipeline over a box
int w8_choose_axis(const Geo& g, const int* box);
// the tiles' partials of dot(z, r) over a box, as the pipelined backward sweep forms them
template <class T>
void w8_dot_tiles(const Geo& g, int ax, const int* box, const T* z, const T* r, double* partial, const PcgScalars* S,
                  hipStream_t st);

// ---------------- particle kernels (xf_particles.hip) ----------------
struct Particles {
    int64_t n, cap;
    double *px, *py, *pz, *vx, *vy, *vz;
    int32_t* pid;
    // sort scratch
    double *tx, *ty, *tz, *tvx, *tvy, *tvz;
    int32_t *tpid, *key, *perm;
};
struct CellIndex {
    int32_t* count;   // N+1 (last bin: out of range)
    int32_t* start;   // N+2
    int32_t* cursor;  // N+1
    int32_t* scan_tmp;
    int scan_tmp_n;
};
void sort_particles(const Geo& g, Particles& P, CellIndex& ci, hipStream_t st);
// n x 3 interleaved device arrays in id order <-> the SoA arrays (from_n3 sets
// the ids to 0..n-1; pos or vel may be null in to_n3)
void particles_to_n3(const Particles& P, double* pos, double* vel, hipStream_t st);
void particles_from_n3(Particles& P, const double* pos, const double* vel, hipStream_t st);
void launch_g2p_advect(const Geo& g, Particles& P, const uint8_t* mat, const double* u, const double* v,
                       const double* w, const double* us, const double* vs, const double* ws, double dt, double flip,
                       int integrator, int64_t limit_id, hipStream_t st);
// one fluid source's update: mark its cells fluid, write ncells * per_cell
// particles at n0.. (the caller has grown P)
void launch_source(const Geo& g, Particles& P, uint8_t* mat, const int32_t* cells, int ncells, int per_cell,
                   const double v[3], uint64_t seed, uint64_t m0, int64_t n0, hipStream_t st);
void launch_p2g(const Geo& g, const Particles& P, const CellIndex& ci, double* f, int axis, int kernel,
                hipStream_t st);
void launch_density(const Geo& g, const Particles& P, const CellIndex& ci, double* phi, int d, double radius,
                    hipStream_t st);
int64_t surface_cells(const Geo& g, const double* phi, int d, double iso, int32_t* cells, uint8_t* cube,
                      int64_t cap, int32_t* count_tmp, int32_t* start_tmp, int32_t* scan_tmp, int scan_tmp_n,
                      hipStream_t st);
void launch_maxv2(const Particles& P, double* partial, int nblocks, hipStream_t st);
void exclusive_scan_i32(const int32_t* in, int32_t* out, long n, int32_t* tmp, int tmp_n, hipStream_t st);
long scan_tmp_size(long n);

// ---------------- surface mesh (xf_mesh.hip) ----------------
// Device buffers of the last marching-cubes mesh (grown on demand).
struct MeshWork {
    int32_t* owner;  // per edge of the density lattice: claiming cell, then vertex index
    long nowner;
    int32_t* cells;
    uint8_t* cube;
    int32_t *flag, *vpos, *tcnt, *tpos, *scan_tmp;
    double *verts, *verts2, *normals, *fnorm;
    int32_t* tris;
    int32_t *vt_cnt, *vt_start, *vt_cursor, *vt_list;
    long cap_cells, cap_cube, cap_flag, cap_vpos, cap_tcnt, cap_tpos, cap_scan, cap_verts, cap_verts2, cap_normals,
        cap_fnorm, cap_tris, cap_vt_cnt, cap_vt_start, cap_vt_cursor, cap_vt_list;
    long ncells, nverts, ntris;
};
// Builds the mesh of phi: 0 ok, -1 lattice too large, -2 allocation failure.
int build_mesh(const Geo& g, const double* phi, int d, double iso, int smooth_iters, double smooth_value,
               int32_t* count_tmp, int32_t* start_tmp, int32_t* scan_tmp, int scan_tmp_n, MeshWork& w,
               hipStream_t st);
void mesh_free(MeshWork& w);

// ---------------- adjoint kernels (xf_adjoint.hip) ----------------
// pbar = -G^T g (transpose of the pressure application over all three face arrays)
void launch_apply_adjoint(const Geo& g, const uint8_t* mat, const double* gu, const double* gv, const double* gw,
                          double* pbar, hipStream_t st);
// f += B^T bbar for one face array (transpose of the right-hand-side assembly)
void launch_rhs_adjoint(const Geo& g, const uint8_t* mat, const double* bbar, double* f, int axis, hipStream_t st);
// particle transfers (particles as n x 3 arrays); grid gradients accumulate
void launch_sample_adjoint(const Geo& g, const double* const F[3], double* const G[3], long n, const double* pos,
                           const double* gvel, double* gpos, hipStream_t st);
// the position update with its collision branch when mat is given
void launch_advect_adjoint(const Geo& g, const double* const F[3], double* const G[3], const uint8_t* mat, long n,
                           const double* pos, double dt, int integrator, const double* gnew, double* gpos,
                           hipStream_t st);
void launch_density_adjoint(const Geo& g, int d, double R, long n, const double* pos, const double* gphi,
                            double* gpos, hipStream_t st);
void launch_p2g_adjoint(const Geo& g, int kernel, long n, const double* pos, const double* vel,
                        const double* const gf[3], double* gvel, double* gpos, double* s_wt, double* s_val,
                        hipStream_t st);
