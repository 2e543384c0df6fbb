// xf_capi.hip -- the C ABI (include/xfluid.h): device state of one simulator
// and the host orchestration of FluidSimulator::frameStep / timeStep
// (FluidSimulator.cpp:224-335) over the gfx950 kernels.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <mutex>
#include "../../include/xfluid.h"
#include "xf_internal.h"
static_assert(OP_VARIATIONAL == XF_PRESSURE_VARIATIONAL && OP_NORMAL == XF_PRESSURE_NORMAL,
              "pressure operator ids");

namespace {
thread_local std::string g_err;
int set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return set_err(-2, "%s: %s", #x, hipGetErrorString(e_));     \
    } while (0)

template <class T>
hipError_t dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}
template <class T>
void dfree(T*& p) {
    if (p) hipFree(p);
    p = nullptr;
}

inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double u01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
}  // namespace

// Scratch of the allocation-free P2G adjoint (xf_sim_p2g_adjoint_noalloc): the
// weight sums and value sums of one face array, XF_P2G_ADJOINT_NOALLOC_MAX_FACES
// doubles each.  One copy per GPU, held until the process ends.  Host code never
// passes this name to a kernel: it uses the per-device address that
// hipGetSymbolAddress returns (scratch_setup, at simulator creation).
static __device__ __align__(256) double scratch[2 * (size_t)XF_P2G_ADJOINT_NOALLOC_MAX_FACES];

namespace {
const int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The device address of `scratch` on device dev, which must be current.
// Looked up once per device; 0 on success.
int scratch_setup(int dev) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return set_err(-1, "device %d out of range for the scratch table", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_err(-2, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)ptr;
    return 0;
}
double* scratch_ptr(int dev) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}
}  // namespace

struct xf_sim {
    Geo g{};
    xf_params prm{};
    int dev = 0;
    hipStream_t st = nullptr;
    std::vector<uint8_t> h_mat;
    std::vector<double> forces;
    double *u = nullptr, *v = nullptr, *w = nullptr, *us = nullptr, *vs = nullptr, *ws = nullptr;
    uint8_t* mat = nullptr;
    uint16_t* code = nullptr;
    int8_t* marker = nullptr;
    double* phi = nullptr;
    long nphi = 0;
    int phi_d = 0;
    Particles P{};
    CellIndex ci{};
    bool sorted = false;
    PcgWork pw{};
    double* red = nullptr;
    double* h_red = nullptr;
    int32_t* surf_cnt = nullptr;
    int32_t* surf_start = nullptr;
    xf_solve_stats last{};
    int frame = 0;
    bool timing = false;
    xf_step_timing tlast{};
    hipEvent_t ev[XF_STAGE_COUNT + 1]{};
    bool inited = false;
    MeshWork mesh{};
    double *ga_u = nullptr, *ga_v = nullptr, *ga_w = nullptr, *ga_p = nullptr;  // adjoint scratch
    bool mesh_out = false;
    std::string mesh_prefix;
    // fluid sources (fluidsource.h:10-53): cells fixed at creation, velocity, |v|
    struct Source {
        int32_t* cells = nullptr;
        int ncells = 0;
        double v[3] = {0, 0, 0};
        double normv = 0;
    };
    std::vector<Source> sources;
    uint64_t source_seed = 0, source_emitted = 0;  // stream seed, particles emitted so far

    long nu() const { return (long)(g.ni + 1) * g.nj * g.nk; }
    long nv() const { return (long)g.ni * (g.nj + 1) * g.nk; }
    long nw() const { return (long)g.ni * g.nj * (g.nk + 1); }
    long maxface() const { return std::max(nu(), std::max(nv(), nw())); }
};

namespace {
const int RED_N = 1024;

long phi_count(const Geo& g, int d) { return (long)(d * g.ni + 1) * (d * g.nj + 1) * (d * g.nk + 1); }

int ensure_phi(xf_sim* s) {
    int d = s->prm.sub_division;
    if (s->phi && s->phi_d == d) return 0;
    dfree(s->phi);
    s->nphi = phi_count(s->g, d);
    HIPCHK(dalloc(&s->phi, s->nphi));
    HIPCHK(hipMemsetAsync(s->phi, 0, s->nphi * 8, s->st));
    s->phi_d = d;
    return 0;
}

int ensure_particles(xf_sim* s, int64_t n) {
    Particles& P = s->P;
    if (n <= P.cap) { P.n = n; return 0; }
    double** d[] = {&P.px, &P.py, &P.pz, &P.vx, &P.vy, &P.vz, &P.tx, &P.ty, &P.tz, &P.tvx, &P.tvy, &P.tvz};
    for (auto* p : d) dfree(*p);
    dfree(P.pid); dfree(P.tpid); dfree(P.key); dfree(P.perm);
    for (auto* p : d) HIPCHK(dalloc(p, n));
    HIPCHK(dalloc(&P.pid, n)); HIPCHK(dalloc(&P.tpid, n)); HIPCHK(dalloc(&P.key, n)); HIPCHK(dalloc(&P.perm, n));
    P.cap = n;
    P.n = n;
    return 0;
}

// grow the particle arrays to hold `need` particles, keeping the first P.n
// (capacity grows by at least half, so a source appending every substep
// reallocates rarely); the sort scratch is reallocated without copying
int grow_particles(xf_sim* s, int64_t need) {
    Particles& P = s->P;
    if (need <= P.cap) return 0;
    const int64_t cap = std::max<int64_t>(need, P.cap + P.cap / 2);
    double** keep[] = {&P.px, &P.py, &P.pz, &P.vx, &P.vy, &P.vz};
    double** scratch[] = {&P.tx, &P.ty, &P.tz, &P.tvx, &P.tvy, &P.tvz};
    int32_t** iscratch[] = {&P.tpid, &P.key, &P.perm};
    double* nk[6] = {};
    int32_t* npid = nullptr;
    bool ok = dalloc(&npid, cap) == hipSuccess;
    for (int a = 0; a < 6; ++a) ok = ok && dalloc(&nk[a], cap) == hipSuccess;
    if (!ok) {
        for (auto& q : nk) dfree(q);
        dfree(npid);
        return set_err(-2, "particle allocation failed (%lld particles)", (long long)cap);
    }
    for (int a = 0; a < 6; ++a)
        if (P.n) HIPCHK(hipMemcpyAsync(nk[a], *keep[a], P.n * 8, hipMemcpyDeviceToDevice, s->st));
    if (P.n) HIPCHK(hipMemcpyAsync(npid, P.pid, P.n * 4, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    for (int a = 0; a < 6; ++a) { dfree(*keep[a]); *keep[a] = nk[a]; }
    dfree(P.pid);
    P.pid = npid;
    for (auto* q : scratch) { dfree(*q); HIPCHK(dalloc(q, cap)); }
    for (auto* q : iscratch) { dfree(*q); HIPCHK(dalloc(q, cap)); }
    P.cap = cap;
    return 0;
}

// _applyFluidSource (FluidSimulator.cpp:1087-1093) at the end of the advection
// stage (FluidSimulator.cpp:725-726): sources in creation order, each appending
// (int)(8 |v| dt / dx) particles per cell (fluidsource.cpp:77)
int apply_sources(xf_sim* s, double dt) {
    if (s->sources.empty()) return 0;
    const Geo& g = s->g;
    std::vector<int> per(s->sources.size());
    int64_t add = 0;
    for (size_t q = 0; q < s->sources.size(); ++q) {
        const double x = 8.0 * s->sources[q].normv * dt / g.dx;  // _particlePerCell * _normv * dt / _dx
        if (!(x < 2147483647.0)) return set_err(-1, "fluid source %d: %g particles per cell", (int)q, x);
        per[q] = x > 0 ? (int)x : 0;
        add += (int64_t)per[q] * s->sources[q].ncells;
    }
    int64_t n = s->P.n;
    if (n + add > 2147483647LL) return set_err(-1, "fluid sources: particle count would pass 2^31 - 1");
    if (grow_particles(s, n + add)) return -2;
    for (size_t q = 0; q < s->sources.size(); ++q) {
        const auto& src = s->sources[q];
        launch_source(g, s->P, s->mat, src.cells, src.ncells, per[q], src.v, s->source_seed, s->source_emitted, n,
                      s->st);
        n += (int64_t)per[q] * src.ncells;
        s->source_emitted += (uint64_t)per[q] * src.ncells;
    }
    HIPCHK(hipGetLastError());
    s->P.n = n;
    s->sorted = false;
    return 0;
}

int upload_mat(xf_sim* s) {
    HIPCHK(hipMemcpyAsync(s->mat, s->h_mat.data(), s->g.N, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int download_mat(xf_sim* s) {
    HIPCHK(hipMemcpyAsync(s->h_mat.data(), s->mat, s->g.N, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}

double particle_radius(const xf_sim* s) {  // FluidSimulator.cpp:380-385
    double dx = s->g.dx;
    double volume = dx * dx * dx / 8.0;
    double pi = 3.141592653;
    return std::pow(3 * volume / (4 * pi), 1.0 / 3.0);
}
double blob_radius(const xf_sim* s) {
    return s->prm.blob_radius > 0 ? s->prm.blob_radius : particle_radius(s) * 3.0;
}

int ensure_sorted(xf_sim* s) {
    if (s->sorted) return 0;
    sort_particles(s->g, s->P, s->ci, s->st);
    HIPCHK(hipGetLastError());
    s->sorted = true;
    return 0;
}

// ---- stages ----
int st_density(xf_sim* s) {
    int rc = ensure_phi(s);
    if (rc) return rc;
    rc = ensure_sorted(s);
    if (rc) return rc;
    launch_density(s->g, s->P, s->ci, s->phi, s->prm.sub_division, blob_radius(s), s->st);
    HIPCHK(hipGetLastError());
    return 0;
}
int st_material(xf_sim* s) {
    int rc = ensure_sorted(s);
    if (rc) return rc;
    launch_material(s->g, s->mat, s->ci.count, s->st);
    HIPCHK(hipGetLastError());
    return 0;
}
int st_save(xf_sim* s) {
    HIPCHK(hipMemcpyAsync(s->us, s->u, s->nu() * 8, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipMemcpyAsync(s->vs, s->v, s->nv() * 8, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipMemcpyAsync(s->ws, s->w, s->nw() * 8, hipMemcpyDeviceToDevice, s->st));
    return 0;
}
int st_forces(xf_sim* s, double dt) {  // FluidSimulator.cpp:429-499
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (size_t f = 0; f < s->forces.size() / 3; ++f) {
        ax += s->forces[3 * f]; ay += s->forces[3 * f + 1]; az += s->forces[3 * f + 2];
    }
    double dv[3] = {ax * dt, ay * dt, az * dt};
    double* fa[3] = {s->u, s->v, s->w};
    const double err = 1e-10;
    for (int a = 0; a < 3; ++a)
        if (dv[a] > err || dv[a] < -err) launch_forces(s->g, s->mat, fa[a], a, dv[a], s->st);
    HIPCHK(hipGetLastError());
    return 0;
}
// Pressure solve of the substep: stencil codes and RHS from the face arrays
// (fu, fv, fw), then PCG into pw.x.  With rhs_override the RHS is replaced by
// rhs_override (device, N values) after the stencil codes are built.
int st_solve_from(xf_sim* s, const double* fu, const double* fv, const double* fw, const double* rhs_override) {
    // (events only with timing on: an event waited for after the solve's own
    // wait costs another round trip to the device)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (s->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, s->st));
    }
    launch_stencil_rhs(s->g, s->mat, fu, fv, fw, s->code, s->pw.b, s->st, s->prm.pressure_solver);
    if (rhs_override)
        HIPCHK(hipMemcpyAsync(s->pw.b, rhs_override, s->g.N * 8, hipMemcpyDeviceToDevice, s->st));
    s->last.fallback = 0;
    const xf_params& p = s->prm;
    SweepTimer& tm = s->pw.timer;
    tm.on = s->timing;
    tm.used = 0;
    if (tm.on && !tm.ev) {
        tm.cap = 4096;
        tm.ev = new hipEvent_t[tm.cap];
        for (int i = 0; i < tm.cap; ++i) HIPCHK(hipEventCreate(&tm.ev[i]));
    }
    int sweep = p.sweep_kernel == 0 ? 9 : p.sweep_kernel;
    s->pw.sk64.fused = s->pw.sk32.fused = -1;
    PcgResult r = pcg_solve(s->g, s->code, s->pw, p.pcg_tol, p.pcg_max_iter, p.mic_tau, p.mic_sigma, p.precision,
                            p.mixed_inner_tol, p.mixed_max_outer, sweep, s->st);
    if (sweep >= 2) {
        int err = r.sweep_err;
        if (!r.err_known) {
            // (a pinned destination: the copy joins the stream, and one wait covers the
            // solve's last kernels and the copy)
            HIPCHK(hipMemcpyAsync(s->pw.err_host, s->pw.sweep_flags, sizeof(int), hipMemcpyDeviceToHost, s->st));
            HIPCHK(hipStreamSynchronize(s->st));
            err = *(volatile int*)s->pw.err_host;
        }
        if (err) {  // a pipelined wait timed out: redo the solve with per-hyperplane launches
            fprintf(stderr, "xfluid: pipelined sweep wait timed out; solving with per-plane sweeps\n");
            HIPCHK(hipMemsetAsync(s->pw.sweep_flags, 0, sizeof(int), s->st));
            tm.used = 0;
            r = pcg_solve(s->g, s->code, s->pw, p.pcg_tol, p.pcg_max_iter, p.mic_tau, p.mic_sigma, p.precision,
                          p.mixed_inner_tol, p.mixed_max_outer, 1, s->st);
            s->last.fallback = 1;
        }
    }
    float ms = 0;
    if (s->timing) {
        HIPCHK(hipEventRecord(e1, s->st));
        HIPCHK(hipEventSynchronize(e1));
        hipEventElapsedTime(&ms, e0, e1);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
    } else if (sweep < 2) {
        HIPCHK(hipStreamSynchronize(s->st));
    }
    HIPCHK(hipGetLastError());
    s->last.status = r.status;
    s->last.iterations = r.iterations;
    s->last.outer = r.outer;
    s->last.residual = r.residual;
    s->last.rhs_norm = r.rhs_norm;
    s->last.box_cells = r.box_cells;
    s->last.sweep_fused = std::max(s->pw.sk64.fused, s->pw.sk32.fused);
    s->last.ms = ms;
    s->last.sweep_ms = 0.0;
    s->last.sweeps = 0;
    for (int i = 0; i + 1 < tm.used; i += 2) {
        float d = 0;
        if (hipEventElapsedTime(&d, tm.ev[i], tm.ev[i + 1]) == hipSuccess) {
            s->last.sweep_ms += d;
            s->last.sweeps += 2;
        }
    }
    tm.used = 0;
    return r.status;
}
int st_solve(xf_sim* s) { return st_solve_from(s, s->u, s->v, s->w, nullptr); }
int st_apply(xf_sim* s) {
    double* fa[3] = {s->u, s->v, s->w};
    launch_apply_pressure(s->g, s->mat, s->pw.x, fa[0], fa[1], fa[2], s->st, s->prm.pressure_solver);
    HIPCHK(hipGetLastError());
    return 0;
}
int st_extrapolate(xf_sim* s) {  // FluidSimulator.cpp:511-519
    int layers = (int)(s->prm.cfl + 5);
    if (layers > 120) return set_err(-1, "extrapolation layers %d exceed 120", layers);
    double* fa[3] = {s->u, s->v, s->w};
    launch_extrapolate(s->g, s->mat, fa[0], fa[1], fa[2], s->marker, s->maxface(), layers, s->st);
    HIPCHK(hipGetLastError());
    return 0;
}
int st_advect(xf_sim* s, double dt) {
    int64_t n = s->P.n;
    int64_t limit = s->prm.tail_skip ? (n / 6) * 6 : n;  // FluidSimulator.cpp:654-663
    launch_g2p_advect(s->g, s->P, s->mat, s->u, s->v, s->w, s->us, s->vs, s->ws, dt, s->prm.flip_ratio,
                      s->prm.integrator, limit, s->st);
    HIPCHK(hipGetLastError());
    s->sorted = false;
    return apply_sources(s, dt);
}
int st_p2g(xf_sim* s) {
    int rc = ensure_sorted(s);
    if (rc) return rc;
    double* fa[3] = {s->u, s->v, s->w};
    for (int a = 0; a < 3; ++a) launch_p2g(s->g, s->P, s->ci, fa[a], a, s->prm.p2g_kernel, s->st);
    HIPCHK(hipGetLastError());
    return 0;
}

int run_stage(xf_sim* s, int stage, double dt) {
    switch (stage) {
        case XF_STAGE_DENSITY: return st_density(s);
        case XF_STAGE_MATERIAL: return st_material(s);
        case XF_STAGE_SAVE: return st_save(s);
        case XF_STAGE_FORCES: return st_forces(s, dt);
        case XF_STAGE_SOLVE: return st_solve(s);
        case XF_STAGE_APPLY: return st_apply(s);
        case XF_STAGE_EXTRAPOLATE: return st_extrapolate(s);
        case XF_STAGE_ADVECT: return st_advect(s, dt);
        case XF_STAGE_P2G: return st_p2g(s);
    }
    return set_err(-1, "unknown stage %d", stage);
}

int check(const xf_sim* s) {
    if (!s) return set_err(-1, "null handle");
    return 0;
}
}  // namespace

extern "C" {

const char* xf_last_error(void) { return g_err.c_str(); }

void xf_params_default(xf_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->flip_ratio = 0.95; p->cfl = 1.0; p->max_dt = 0.5; p->pcg_tol = 1e-5; p->pcg_max_iter = 100;
    p->sub_division = 2; p->mic_tau = 0.97; p->mic_sigma = 0.25; p->blob_radius = 0.0; p->iso = 0.4;
    p->precision = XF_PREC_FP64; p->integrator = XF_INTEGRATOR_RK4; p->p2g_kernel = XF_P2G_WYVILL;
    p->tail_skip = 1; p->density_every_substep = 1; p->sweep_kernel = 0;
    p->mixed_inner_tol = 1e-3; p->mixed_max_outer = 20;
}

int xf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int xf_sim_create(int ni, int nj, int nk, xf_sim** out) {
    if (!out) return set_err(-1, "null out");
    *out = nullptr;
    if (ni < 3 || nj < 3 || nk < 3) return set_err(-1, "grid must be at least 3x3x3 (got %d,%d,%d)", ni, nj, nk);
    if ((long)ni * nj * nk >= (1L << 30)) return set_err(-1, "grid too large");
    if (xf_device_count() < 1) return set_err(-3, "no HIP device: the xfluid HIP library needs an MI355X (gfx950)");
    xf_sim* s = new xf_sim();
    s->g.ni = ni; s->g.nj = nj; s->g.nk = nk;
    s->g.nn = (long)ni * nj;
    s->g.N = s->g.nn * nk;
    s->g.dx = 0.0001;  // FluidSimulator.h:175
    xf_params_default(&s->prm);
    hipGetDevice(&s->dev);
    auto fail = [&](int rc) { xf_sim_destroy(s); return rc; };
    if (scratch_setup(s->dev)) return fail(-2);
    if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) return fail(set_err(-2, "stream"));
    const long N = s->g.N;
    bool ok = dalloc(&s->u, s->nu()) == hipSuccess && dalloc(&s->v, s->nv()) == hipSuccess &&
              dalloc(&s->w, s->nw()) == hipSuccess && dalloc(&s->us, s->nu()) == hipSuccess &&
              dalloc(&s->vs, s->nv()) == hipSuccess && dalloc(&s->ws, s->nw()) == hipSuccess &&
              dalloc(&s->mat, N) == hipSuccess && dalloc(&s->code, N) == hipSuccess &&
              dalloc(&s->marker, 3 * s->maxface()) == hipSuccess && dalloc(&s->ci.count, N + 1) == hipSuccess &&
              dalloc(&s->ci.start, N + 2) == hipSuccess && dalloc(&s->ci.cursor, N + 1) == hipSuccess &&
              dalloc(&s->pw.x, N) == hipSuccess && dalloc(&s->pw.r, N) == hipSuccess &&
              dalloc(&s->pw.s, N) == hipSuccess && dalloc(&s->pw.s2, N) == hipSuccess &&
              dalloc(&s->pw.z, N) == hipSuccess && dalloc(&s->pw.s2_32, N) == hipSuccess &&
              dalloc(&s->pw.q, N) == hipSuccess && dalloc(&s->pw.b, N) == hipSuccess &&
              dalloc(&s->pw.inv, N) == hipSuccess && dalloc(&s->pw.x32, N) == hipSuccess &&
              dalloc(&s->pw.r32, N) == hipSuccess && dalloc(&s->pw.s32, N) == hipSuccess &&
              dalloc(&s->pw.z32, N) == hipSuccess && dalloc(&s->pw.q32, N) == hipSuccess &&
              dalloc(&s->pw.inv32, N) == hipSuccess &&
              dalloc(&s->pw.S, 2) == hipSuccess && dalloc(&s->red, RED_N) == hipSuccess &&
              dalloc(&s->surf_cnt, N + 1) == hipSuccess && dalloc(&s->surf_start, N + 2) == hipSuccess;
    if (!ok) return fail(set_err(-2, "device allocation failed"));
    s->pw.nflags = sweep_flag_words(s->g);
    s->pw.epoch = 1;
    if (dalloc(&s->pw.sweep_flags, s->pw.nflags) != hipSuccess) return fail(set_err(-2, "alloc flags"));
    if (dalloc(&s->pw.box_dev, 8) != hipSuccess || hipHostMalloc((void**)&s->pw.box_host, 8 * sizeof(int)) != hipSuccess)
        return fail(set_err(-2, "alloc box"));
    {
        SweepSkew<double>& a = s->pw.sk64;
        SweepSkew<float>& b = s->pw.sk32;
        const long W8 = sweep_w8_size(s->g), E8 = sweep_w8_edge_words(s->g);
        if (dalloc(&a.oRec, 8 * W8) != hipSuccess || dalloc(&a.oQ, 2 * W8) != hipSuccess ||
            dalloc(&a.oEdge, E8) != hipSuccess || dalloc(&b.oRec, 8 * W8) != hipSuccess ||
            dalloc(&b.oQ, 2 * W8) != hipSuccess || dalloc(&b.oEdge, E8) != hipSuccess)
            return fail(set_err(-2, "alloc sweep buffers"));
        // tags are never zero: zeroed edge words read as not yet written
        hipMemsetAsync(a.oEdge, 0, E8 * sizeof(uint32_t), s->st);
        hipMemsetAsync(b.oEdge, 0, E8 * sizeof(uint32_t), s->st);
    }
    hipMemsetAsync(s->pw.sweep_flags, 0, s->pw.nflags * sizeof(int), s->st);
    s->ci.scan_tmp_n = (int)scan_tmp_size(N + 2);
    if (dalloc(&s->ci.scan_tmp, s->ci.scan_tmp_n) != hipSuccess) return fail(set_err(-2, "alloc scan"));
    {
        // partial halves: the reduction blocks, or the tiles of a whole-grid box
        const int all[6] = {0, s->g.ni - 1, 0, s->g.nj - 1, 0, s->g.nk - 1};
        const int t0 = w8_box_tiles(s->g, 0, all), t2 = w8_box_tiles(s->g, 2, all);
        const int tiles = t0 > t2 ? t0 : t2;
        s->pw.npartial = tiles > RED_N ? tiles : RED_N;
        if (dalloc(&s->pw.partial, 2 * (long)s->pw.npartial) != hipSuccess) return fail(set_err(-2, "alloc partials"));
    }
    if (hipHostMalloc((void**)&s->pw.S_host, 3 * sizeof(PcgScalars)) != hipSuccess ||
        hipEventCreateWithFlags(&s->pw.chunk_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s->pw.chunk_ev[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&s->h_red, RED_N * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void**)&s->pw.err_host, sizeof(int)) != hipSuccess)
        return fail(set_err(-2, "pinned alloc"));
    std::memset(s->pw.S_host, 0, 3 * sizeof(PcgScalars));
    hipMemsetAsync(s->u, 0, s->nu() * 8, s->st);
    hipMemsetAsync(s->v, 0, s->nv() * 8, s->st);
    hipMemsetAsync(s->w, 0, s->nw() * 8, s->st);
    hipMemsetAsync(s->us, 0, s->nu() * 8, s->st);
    hipMemsetAsync(s->vs, 0, s->nv() * 8, s->st);
    hipMemsetAsync(s->ws, 0, s->nw() * 8, s->st);
    hipMemsetAsync(s->pw.x, 0, N * 8, s->st);
    hipMemsetAsync(s->pw.b, 0, N * 8, s->st);
    hipMemsetAsync(s->pw.inv, 0, N * 8, s->st);
    hipMemsetAsync(s->code, 0, N * 2, s->st);
    s->h_mat.assign(N, 0);
    if (upload_mat(s)) return fail(-2);
    if (ensure_phi(s)) return fail(-2);
    for (auto& e : s->ev) hipEventCreate(&e);
    if (hipStreamSynchronize(s->st) != hipSuccess) return fail(set_err(-2, "init sync"));
    *out = s;
    return 0;
}

void xf_sim_destroy(xf_sim* s) {
    if (!s) return;
    if (s->st) hipStreamSynchronize(s->st);
    double** ds[] = {&s->u, &s->v, &s->w, &s->us, &s->vs, &s->ws, &s->phi, &s->pw.x, &s->pw.r, &s->pw.s, &s->pw.s2, &s->pw.z,
                     &s->pw.q, &s->pw.b, &s->pw.inv, &s->pw.partial, &s->red, &s->P.px, &s->P.py, &s->P.pz,
                     &s->P.vx, &s->P.vy, &s->P.vz, &s->P.tx, &s->P.ty, &s->P.tz, &s->P.tvx, &s->P.tvy, &s->P.tvz};
    for (auto* p : ds) dfree(*p);
    float** fs[] = {&s->pw.x32, &s->pw.r32, &s->pw.s32, &s->pw.s2_32, &s->pw.z32, &s->pw.q32, &s->pw.inv32};
    for (auto* p : fs) dfree(*p);
    int32_t** is[] = {&s->ci.count, &s->ci.start, &s->ci.cursor, &s->ci.scan_tmp, &s->P.pid, &s->P.tpid,
                      &s->P.key, &s->P.perm, &s->surf_cnt, &s->surf_start};
    for (auto* p : is) dfree(*p);
    dfree(s->mat); dfree(s->code); dfree(s->marker); dfree(s->pw.S); dfree(s->pw.sweep_flags);
    dfree(s->pw.box_dev);
    if (s->pw.box_host) { hipHostFree(s->pw.box_host); s->pw.box_host = nullptr; }
    dfree(s->pw.sk64.oRec); dfree(s->pw.sk64.oQ); dfree(s->pw.sk64.oEdge);
    dfree(s->pw.sk32.oRec); dfree(s->pw.sk32.oQ); dfree(s->pw.sk32.oEdge);
    mesh_free(s->mesh);
    for (auto& src : s->sources) dfree(src.cells);
    dfree(s->ga_u); dfree(s->ga_v); dfree(s->ga_w); dfree(s->ga_p);
    if (s->pw.timer.ev) {
        for (int i = 0; i < s->pw.timer.cap; ++i) hipEventDestroy(s->pw.timer.ev[i]);
        delete[] s->pw.timer.ev;
    }
    if (s->pw.S_host) hipHostFree(s->pw.S_host);
    if (s->pw.err_host) hipHostFree(s->pw.err_host);
    for (hipEvent_t e : s->pw.chunk_ev)
        if (e) hipEventDestroy(e);
    if (s->h_red) hipHostFree(s->h_red);
    for (auto& e : s->ev) if (e) hipEventDestroy(e);
    if (s->st) hipStreamDestroy(s->st);
    delete s;
}

int xf_sim_set_cell_width(xf_sim* s, double dx) {
    if (check(s)) return -1;
    if (!(dx > 0)) return set_err(-1, "dx must be positive");
    s->g.dx = dx;
    return 0;
}
int xf_sim_set_params(xf_sim* s, const xf_params* p) {
    if (check(s) || !p) return set_err(-1, "null argument");
    if (p->sub_division < 1 || p->sub_division > 4) return set_err(-1, "sub_division must be 1..4");
    if (p->pcg_max_iter < 0) return set_err(-1, "pcg_max_iter < 0");
    if (p->sweep_kernel != 0 && p->sweep_kernel != 1 && p->sweep_kernel != 9)
        return set_err(-1, "sweep_kernel must be 0 (auto), 1 (per-hyperplane) or 9 (8x8 line tiles), got %d",
                       p->sweep_kernel);
    if (p->pressure_solver != XF_PRESSURE_VARIATIONAL && p->pressure_solver != XF_PRESSURE_NORMAL)
        return set_err(-1, "unknown pressure_solver %d", p->pressure_solver);
    s->prm = *p;
    return ensure_phi(s);
}
int xf_sim_get_params(const xf_sim* s, xf_params* p) {
    if (check(s) || !p) return set_err(-1, "null argument");
    *p = s->prm;
    return 0;
}

// FluidSimulator.cpp:64-96
int xf_sim_set_solid_boundary(xf_sim* s) {
    if (check(s)) return -1;
    if (download_mat(s)) return -2;
    const Geo& g = s->g;
    for (int k = 0; k < g.nk; ++k)
        for (int j = 0; j < g.nj; ++j)
            for (int i = 0; i < g.ni; ++i)
                if (i == 0 || j == 0 || k == 0 || i == g.ni - 1 || j == g.nj - 1 || k == g.nk - 1)
                    s->h_mat[cidx(g, i, j, k)] = XF_SOLID;
    return upload_mat(s);
}
// FluidSimulator.cpp:103-121 (r2 held in float as in the reference)
int xf_sim_add_sphere_fluid(xf_sim* s, double x, double y, double z, double r) {
    if (check(s)) return -1;
    if (download_mat(s)) return -2;
    const Geo& g = s->g;
    float r2 = r * r;
    for (int i = 1; i < g.ni - 1; ++i)
        for (int j = 1; j < g.nj - 1; ++j)
            for (int k = 1; k < g.nk - 1; ++k) {
                double ex = x - (i + 0.5) * g.dx, ey = y - (j + 0.5) * g.dx, ez = z - (k + 0.5) * g.dx;
                if (ex * ex + ey * ey + ez * ez <= r2) s->h_mat[cidx(g, i, j, k)] = XF_FLUID;
            }
    return upload_mat(s);
}
// FluidSimulator.cpp:123-140
int xf_sim_add_cubic_fluid(xf_sim* s, const double lu[3], const double rd[3]) {
    if (check(s) || !lu || !rd) return set_err(-1, "null argument");
    if (download_mat(s)) return -2;
    const Geo& g = s->g;
    for (int i = 1; i < g.ni - 1; ++i)
        for (int j = 1; j < g.nj - 1; ++j)
            for (int k = 1; k < g.nk - 1; ++k) {
                double gx = (i + 0.5) * g.dx, gy = (j + 0.5) * g.dx, gz = (k + 0.5) * g.dx;
                if ((gx <= rd[0] && gx >= lu[0]) && (gy <= rd[1] && gy >= lu[1]) && (gz <= rd[2] && gz >= lu[2]))
                    s->h_mat[cidx(g, i, j, k)] = XF_FLUID;
            }
    return upload_mat(s);
}
int xf_sim_add_constant_force(xf_sim* s, double fx, double fy, double fz) {
    if (check(s)) return -1;
    s->forces.push_back(fx); s->forces.push_back(fy); s->forces.push_back(fz);
    return 0;
}

// FluidSimulator::addFluidSource (FluidSimulator.cpp:152-166) with
// CubicFluidSource::setCubic / setVelocity (fluidsource.cpp:8-39)
int xf_sim_add_fluid_source(xf_sim* s, const double c[3], double wx, double wy, double wz, const double vel[3]) {
    if (check(s) || !c || !vel) return set_err(-1, "null argument");
    if (download_mat(s)) return -2;
    const Geo& g = s->g;
    const double lo[3] = {c[0] - wx, c[1] - wy, c[2] - wz}, hi[3] = {c[0] + wx, c[1] + wy, c[2] + wz};
    int a[3], b[3];
    for (int d = 0; d < 3; ++d) {
        const double l = lo[d] / g.dx, h = hi[d] / g.dx;
        if (!(l > -2147483648.0 && l < 2147483647.0 && h > -2147483648.0 && h < 2147483647.0))
            return set_err(-1, "fluid source box outside the index range");
        a[d] = (int)l;  // pos2Index truncates (xfluidutil.cpp:3-6)
        b[d] = (int)h;
    }
    std::vector<int32_t> cells;
    for (int i = std::max(a[0], 0); i <= b[0] && i < g.ni; ++i)
        for (int j = std::max(a[1], 0); j <= b[1] && j < g.nj; ++j)
            for (int k = std::max(a[2], 0); k <= b[2] && k < g.nk; ++k)
                if (s->h_mat[cidx(g, i, j, k)] != XF_SOLID) cells.push_back((int32_t)cidx(g, i, j, k));
    xf_sim::Source src;
    src.v[0] = vel[0]; src.v[1] = vel[1]; src.v[2] = vel[2];
    src.normv = std::sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);  // Vec3::norm
    src.ncells = (int)cells.size();
    if (src.ncells) {
        HIPCHK(dalloc(&src.cells, cells.size()));
        HIPCHK(hipMemcpyAsync(src.cells, cells.data(), cells.size() * 4, hipMemcpyHostToDevice, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    }
    s->sources.push_back(src);
    return 0;
}
int xf_sim_set_source_seed(xf_sim* s, uint64_t seed) {
    if (check(s)) return -1;
    s->source_seed = seed;
    s->source_emitted = 0;
    return 0;
}

static int upload_particles(xf_sim* s, int64_t n, const double* px, const double* py, const double* pz,
                            const double* vx, const double* vy, const double* vz) {
    if (n < 0 || n >= (1LL << 31)) return set_err(-1, "particle count out of range");
    int rc = ensure_particles(s, n);
    if (rc) return rc;
    Particles& P = s->P;
    const double* src[6] = {px, py, pz, vx, vy, vz};
    double* dst[6] = {P.px, P.py, P.pz, P.vx, P.vy, P.vz};
    for (int a = 0; a < 6; ++a)
        if (n) HIPCHK(hipMemcpyAsync(dst[a], src[a], n * 8, hipMemcpyHostToDevice, s->st));
    std::vector<int32_t> ids(n);
    for (int64_t i = 0; i < n; ++i) ids[i] = (int32_t)i;
    if (n) HIPCHK(hipMemcpyAsync(P.pid, ids.data(), n * 4, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    s->sorted = false;
    return 0;
}

// FluidSimulator.cpp:168-184, 347-370 with SplitMix64 jitter
int xf_sim_seed_particles(xf_sim* s, int mode, uint64_t seed) {
    if (check(s)) return -1;
    if (mode != XF_SEED_REFERENCE && mode != XF_SEED_SUBCELL) return set_err(-1, "unknown seed mode");
    if (download_mat(s)) return -2;
    const Geo& g = s->g;
    const double dx = g.dx, q = 0.25 * dx;
    const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, -1, 1}, {-1, -1, 1},
                          {-1, 1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, 1}};
    std::vector<double> P[3];
    for (int i = 0; i < g.ni; ++i)
        for (int j = 0; j < g.nj; ++j)
            for (int k = 0; k < g.nk; ++k) {
                if (s->h_mat[cidx(g, i, j, k)] != XF_FLUID) continue;
                double c[3] = {(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx};
                for (int t = 0; t < 8; ++t)
                    for (int a = 0; a < 3; ++a) {
                        double r = u01(seed);
                        double v;
                        if (mode == XF_SEED_REFERENCE) {
                            double pt = sg[t][a] < 0 ? c[a] - q : c[a] + q;
                            v = pt + (-0.01 + r * (0.01 - -0.01));
                        } else {
                            double lo = sg[t][a] < 0 ? c[a] - 0.5 * dx : c[a];
                            v = lo + r * (0.5 * dx);
                        }
                        P[a].push_back(v);
                    }
            }
    int64_t n = (int64_t)P[0].size();
    std::vector<double> zero(n, 0.0);
    return upload_particles(s, n, P[0].data(), P[1].data(), P[2].data(), zero.data(), zero.data(), zero.data());
}

// FluidSimulator::init (FluidSimulator.cpp:186-222): walls, then particles
// (seeded from the material when none were given).
int xf_sim_init(xf_sim* s) {
    if (check(s)) return -1;
    int rc = xf_sim_set_solid_boundary(s);
    if (rc) return rc;
    if (s->P.n == 0) {
        rc = xf_sim_seed_particles(s, XF_SEED_REFERENCE, 0);
        if (rc) return rc;
    }
    rc = ensure_phi(s);
    s->inited = true;
    return rc;
}

int xf_sim_stage(xf_sim* s, int stage, double dt) {
    if (check(s)) return -1;
    int rc = run_stage(s, stage, dt);
    if (rc < 0) return rc;
    hipError_t e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) return set_err(-2, "stage %d: %s", stage, hipGetErrorString(e));
    return rc;
}

// PLY writer: TriangleMesh::writeMeshToPLY (trianglemesh.cpp:95-252).  With no
// vertices the reference's colour test (vertices.size() == vertexcolors.size())
// holds and it writes the header with colour properties; reproduced.
static int write_ply(const char* path, const std::vector<double>& verts, const std::vector<int32_t>& tris) {
    const size_t nv = verts.size() / 3, nt = tris.size() / 3;
    std::string h = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(nv) +
                    "\nproperty float x\nproperty float y\nproperty float z\n";
    if (nv == 0) h += "property uchar red\nproperty uchar green\nproperty uchar blue\n";
    h += "element face " + std::to_string(nt) + "\nproperty list uchar int vertex_index\nend_header\n";
    std::vector<char> bin(h.begin(), h.end());
    const size_t off = bin.size();
    bin.resize(off + nv * 12 + nt * 13);
    char* q = bin.data() + off;
    for (size_t i = 0; i < 3 * nv; ++i) {
        const float f = (float)verts[i];
        std::memcpy(q, &f, 4);
        q += 4;
    }
    for (size_t t = 0; t < nt; ++t) {
        *q++ = 3;
        std::memcpy(q, &tris[3 * t], 12);
        q += 12;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_err(-1, "cannot open %s", path);
    const size_t w = std::fwrite(bin.data(), 1, bin.size(), f);
    std::fclose(f);
    return w == bin.size() ? 0 : set_err(-2, "short write to %s", path);
}

static int mesh_build(xf_sim* s, int smooth) {
    if (ensure_phi(s)) return -2;
    int rc = build_mesh(s->g, s->phi, s->prm.sub_division, s->prm.iso, smooth ? 4 : 0, 0.8, s->surf_cnt,
                        s->surf_start, s->ci.scan_tmp, s->ci.scan_tmp_n, s->mesh, s->st);
    if (rc == -1) return set_err(-1, "density lattice too large for the mesh edge table");
    if (rc < 0) return set_err(-2, "mesh allocation failed");
    HIPCHK(hipGetLastError());
    return 0;
}
static int mesh_download(xf_sim* s, std::vector<double>* verts, std::vector<int32_t>* tris, std::vector<double>* normals) {
    const MeshWork& m = s->mesh;
    if (verts) {
        verts->resize(3 * m.nverts);
        if (m.nverts) HIPCHK(hipMemcpyAsync(verts->data(), m.verts, 24 * m.nverts, hipMemcpyDeviceToHost, s->st));
    }
    if (tris) {
        tris->resize(3 * m.ntris);
        if (m.ntris) HIPCHK(hipMemcpyAsync(tris->data(), m.tris, 12 * m.ntris, hipMemcpyDeviceToHost, s->st));
    }
    if (normals) {
        normals->resize(3 * m.nverts);
        if (m.nverts) HIPCHK(hipMemcpyAsync(normals->data(), m.normals, 24 * m.nverts, hipMemcpyDeviceToHost, s->st));
    }
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
static int mesh_stage(xf_sim* s) {
    int rc = mesh_build(s, 1);
    if (rc < 0 || !s->mesh_out) return rc;
    std::vector<double> v;
    std::vector<int32_t> t;
    if ((rc = mesh_download(s, &v, &t, nullptr)) < 0) return rc;
    const std::string path = s->mesh_prefix + std::to_string(s->frame) + ".ply";
    // like the reference (an unopenable file writes nothing), a failed frame
    // output does not stop the simulation; it is reported on stderr
    if (write_ply(path.c_str(), v, t) < 0) std::fprintf(stderr, "xfluid: %s\n", xf_last_error());
    return 0;
}

// FluidSimulator::timeStep (FluidSimulator.cpp:265-335)
int xf_sim_time_step(xf_sim* s, double dt, int output_mesh) {
    if (check(s)) return -1;
    const int order[] = {XF_STAGE_DENSITY, XF_STAGE_MATERIAL, XF_STAGE_SAVE, XF_STAGE_FORCES, XF_STAGE_SOLVE,
                         XF_STAGE_APPLY, XF_STAGE_EXTRAPOLATE, XF_STAGE_ADVECT, XF_STAGE_P2G};
    int status = 0;
    if (s->timing) hipEventRecord(s->ev[XF_STAGE_COUNT], s->st);
    hipEvent_t prev = s->ev[XF_STAGE_COUNT];
    std::vector<std::pair<int, hipEvent_t>> marks;
    for (int stg : order) {
        if (stg == XF_STAGE_DENSITY && !(s->prm.density_every_substep || output_mesh)) continue;
        if (stg == XF_STAGE_ADVECT) {
            // sort + P2G below need the positions after advection
        }
        int rc = run_stage(s, stg, dt);
        if (rc < 0) return rc;
        if (rc > 0) status = rc;
        if (stg == XF_STAGE_SAVE && output_mesh) {  // FluidSimulator.cpp:290-305
            rc = mesh_stage(s);
            if (rc < 0) return rc;
        }
        if (s->timing) {
            hipEventRecord(s->ev[stg], s->st);
            marks.push_back({stg, s->ev[stg]});
        }
    }
    hipError_t e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) return set_err(-2, "time step: %s", hipGetErrorString(e));
    if (s->timing) {
        std::memset(&s->tlast, 0, sizeof(s->tlast));
        for (auto& m : marks) {
            float ms = 0;
            hipEventElapsedTime(&ms, prev, m.second);
            s->tlast.ms[m.first] = ms;
            prev = m.second;
        }
        float tot = 0;
        hipEventElapsedTime(&tot, s->ev[XF_STAGE_COUNT], prev);
        s->tlast.total_ms = tot;
    }
    return status;
}

// FluidSimulator::frameStep (FluidSimulator.cpp:224-262) + _maxVelocity (:387-400)
int xf_sim_frame_step(xf_sim* s, double dt) {
    if (check(s)) return -1;
    double left = dt;
    int worst = 0;
    while (left > 0.0) {
        double maxv = 0.0;
        if (s->P.n > 0) {
            launch_maxv2(s->P, s->red, RED_N, s->st);
            HIPCHK(hipMemcpyAsync(s->h_red, s->red, RED_N * 8, hipMemcpyDeviceToHost, s->st));
            HIPCHK(hipStreamSynchronize(s->st));
            for (int i = 0; i < RED_N; ++i) maxv = s->h_red[i] > maxv ? s->h_red[i] : maxv;
        }
        maxv = std::sqrt(maxv);
        double cur = maxv == 0.0 ? s->prm.max_dt : s->prm.cfl * s->g.dx / maxv;
        cur = cur < s->prm.max_dt ? cur : s->prm.max_dt;
        if (left <= cur) { cur = left; left = 0.0; } else left -= cur;
        int rc = xf_sim_time_step(s, cur, left <= 0.0);
        if (rc < 0) return rc;
        if (rc > worst) worst = rc;
    }
    ++s->frame;
    return worst;
}

int xf_sim_last_solve(const xf_sim* s, xf_solve_stats* st) {
    if (check(s) || !st) return set_err(-1, "null argument");
    *st = s->last;
    return 0;
}
int xf_sim_set_timing(xf_sim* s, int on) {
    if (check(s)) return -1;
    s->timing = on != 0;
    return 0;
}
int xf_sim_last_timing(const xf_sim* s, xf_step_timing* t) {
    if (check(s) || !t) return set_err(-1, "null argument");
    *t = s->tlast;
    return 0;
}
int xf_sim_synchronize(xf_sim* s) {
    if (check(s)) return -1;
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int xf_sim_debug_fail_sweeps(xf_sim* s) {
    if (check(s)) return -1;
    // (byte 0 of the error word: the value 1)
    HIPCHK(hipMemsetAsync(s->pw.sweep_flags, 1, 1, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int xf_sim_frame_count(const xf_sim* s) { return s ? s->frame : -1; }

int64_t xf_sim_num_particles(const xf_sim* s) { return s ? s->P.n : -1; }

int xf_sim_set_particles(xf_sim* s, int64_t n, const double* px, const double* py, const double* pz,
                         const double* vx, const double* vy, const double* vz) {
    if (check(s)) return -1;
    if (n > 0 && (!px || !py || !pz || !vx || !vy || !vz)) return set_err(-1, "null particle buffer");
    return upload_particles(s, n, px, py, pz, vx, vy, vz);
}

int xf_sim_set_particles_device(xf_sim* s, int64_t n, const double* pos, const double* vel) {
    if (check(s)) return -1;
    if (n < 0 || n >= (1LL << 31)) return set_err(-1, "particle count out of range");
    if (n > 0 && (!pos || !vel)) return set_err(-1, "null particle buffer");
    int rc = ensure_particles(s, n);
    if (rc) return rc;
    particles_from_n3(s->P, pos, vel, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->st));
    s->sorted = false;
    return 0;
}
int xf_sim_get_particles_device(xf_sim* s, double* pos, double* vel) {
    if (check(s)) return -1;
    particles_to_n3(s->P, pos, vel, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}

// Particles come back in their original order (the device keeps them sorted by cell).
int xf_sim_get_particles(xf_sim* s, double* px, double* py, double* pz, double* vx, double* vy, double* vz) {
    if (check(s)) return -1;
    const int64_t n = s->P.n;
    if (n == 0) return 0;
    std::vector<double> buf(n);
    std::vector<int32_t> ids(n);
    HIPCHK(hipMemcpyAsync(ids.data(), s->P.pid, n * 4, hipMemcpyDeviceToHost, s->st));
    double* src[6] = {s->P.px, s->P.py, s->P.pz, s->P.vx, s->P.vy, s->P.vz};
    double* dst[6] = {px, py, pz, vx, vy, vz};
    HIPCHK(hipStreamSynchronize(s->st));
    for (int a = 0; a < 6; ++a) {
        if (!dst[a]) continue;
        HIPCHK(hipMemcpyAsync(buf.data(), src[a], n * 8, hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        for (int64_t i = 0; i < n; ++i) dst[a][ids[i]] = buf[i];
    }
    return 0;
}

static void* field_dev(xf_sim* s, int which, size_t* bytes) {
    const long N = s->g.N;
    switch (which) {
        case XF_FIELD_U: *bytes = s->nu() * 8; return s->u;
        case XF_FIELD_V: *bytes = s->nv() * 8; return s->v;
        case XF_FIELD_W: *bytes = s->nw() * 8; return s->w;
        case XF_FIELD_USAVED: *bytes = s->nu() * 8; return s->us;
        case XF_FIELD_VSAVED: *bytes = s->nv() * 8; return s->vs;
        case XF_FIELD_WSAVED: *bytes = s->nw() * 8; return s->ws;
        case XF_FIELD_MATERIAL: *bytes = N; return s->mat;
        case XF_FIELD_PHI: ensure_phi(s); *bytes = s->nphi * 8; return s->phi;
        case XF_FIELD_PRESSURE: *bytes = N * 8; return s->pw.x;
        case XF_FIELD_RHS: *bytes = N * 8; return s->pw.b;
        case XF_FIELD_MIC_INV: *bytes = N * 8; return s->pw.inv;
    }
    *bytes = 0;
    return nullptr;
}
size_t xf_sim_field_bytes(const xf_sim* s, int which) {
    size_t b = 0;
    if (!s) return 0;
    field_dev(const_cast<xf_sim*>(s), which, &b);
    return b;
}
int xf_sim_set_field(xf_sim* s, int which, const void* host, size_t bytes) {
    if (check(s) || !host) return set_err(-1, "null argument");
    size_t b = 0;
    void* d = field_dev(s, which, &b);
    if (!d) return set_err(-1, "unknown field %d", which);
    if (b != bytes) return set_err(-1, "field %d holds %zu bytes, got %zu", which, b, bytes);
    HIPCHK(hipMemcpyAsync(d, host, b, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    if (which == XF_FIELD_MATERIAL) std::memcpy(s->h_mat.data(), host, b);
    return 0;
}
int xf_sim_set_field_device(xf_sim* s, int which, const void* dev, size_t bytes) {
    if (check(s) || !dev) return set_err(-1, "null argument");
    size_t b = 0;
    void* d = field_dev(s, which, &b);
    if (!d) return set_err(-1, "unknown field %d", which);
    if (b != bytes) return set_err(-1, "field %d holds %zu bytes, got %zu", which, b, bytes);
    HIPCHK(hipMemcpyAsync(d, dev, b, hipMemcpyDeviceToDevice, s->st));
    if (which == XF_FIELD_MATERIAL) {  // the host copy of the material follows
        HIPCHK(hipMemcpyAsync(s->h_mat.data(), d, b, hipMemcpyDeviceToHost, s->st));
    }
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int xf_sim_get_field_device(xf_sim* s, int which, void* dev, size_t bytes) {
    if (check(s) || !dev) return set_err(-1, "null argument");
    size_t b = 0;
    void* d = field_dev(s, which, &b);
    if (!d) return set_err(-1, "unknown field %d", which);
    if (b != bytes) return set_err(-1, "field %d holds %zu bytes, got %zu", which, b, bytes);
    HIPCHK(hipMemcpyAsync(dev, d, b, hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int xf_sim_get_field(xf_sim* s, int which, void* host, size_t bytes) {
    if (check(s) || !host) return set_err(-1, "null argument");
    size_t b = 0;
    void* d = field_dev(s, which, &b);
    if (!d) return set_err(-1, "unknown field %d", which);
    if (b != bytes) return set_err(-1, "field %d holds %zu bytes, got %zu", which, b, bytes);
    HIPCHK(hipMemcpyAsync(host, d, b, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}

int64_t xf_sim_surface_cells(xf_sim* s, int32_t* cells, uint8_t* cube_index, int64_t cap) {
    if (check(s)) return -1;
    if (ensure_phi(s)) return -2;
    int32_t* dc = nullptr;
    uint8_t* dq = nullptr;
    int64_t total = surface_cells(s->g, s->phi, s->prm.sub_division, s->prm.iso, nullptr, nullptr, 0, s->surf_cnt,
                                  s->surf_start, s->ci.scan_tmp, s->ci.scan_tmp_n, s->st);
    if (cells && cube_index && cap > 0 && total > 0) {
        int64_t m = std::min(cap, total);
        if (dalloc(&dc, m) != hipSuccess || dalloc(&dq, m) != hipSuccess) {
            dfree(dc); dfree(dq);
            return set_err(-2, "surface alloc");
        }
        surface_cells(s->g, s->phi, s->prm.sub_division, s->prm.iso, dc, dq, m, s->surf_cnt, s->surf_start,
                      s->ci.scan_tmp, s->ci.scan_tmp_n, s->st);
        hipMemcpyAsync(cells, dc, m * 4, hipMemcpyDeviceToHost, s->st);
        hipMemcpyAsync(cube_index, dq, m, hipMemcpyDeviceToHost, s->st);
        hipStreamSynchronize(s->st);
        dfree(dc); dfree(dq);
    }
    HIPCHK(hipGetLastError());
    return total;
}

int xf_sim_generate_mesh(xf_sim* s, int smooth, int64_t* nverts, int64_t* ntris) {
    if (check(s)) return -1;
    int rc = mesh_build(s, smooth);
    if (rc < 0) return rc;
    if (nverts) *nverts = s->mesh.nverts;
    if (ntris) *ntris = s->mesh.ntris;
    return 0;
}
int xf_sim_get_mesh(xf_sim* s, double* verts, int32_t* tris, double* normals) {
    if (check(s)) return -1;
    const MeshWork& m = s->mesh;
    if (verts && m.nverts) HIPCHK(hipMemcpyAsync(verts, m.verts, 24 * m.nverts, hipMemcpyDeviceToHost, s->st));
    if (tris && m.ntris) HIPCHK(hipMemcpyAsync(tris, m.tris, 12 * m.ntris, hipMemcpyDeviceToHost, s->st));
    if (normals && m.nverts) HIPCHK(hipMemcpyAsync(normals, m.normals, 24 * m.nverts, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int xf_sim_write_mesh_ply(xf_sim* s, const char* path) {
    if (check(s) || !path) return set_err(-1, "null argument");
    std::vector<double> v;
    std::vector<int32_t> t;
    int rc = mesh_download(s, &v, &t, nullptr);
    if (rc < 0) return rc;
    return write_ply(path, v, t);
}
int xf_sim_set_mesh_output(xf_sim* s, const char* prefix) {
    if (check(s)) return -1;
    s->mesh_out = prefix != nullptr;
    s->mesh_prefix = prefix ? prefix : "";
    return 0;
}

// ---- projection and its adjoint on caller arrays ----
static int copy_faces(xf_sim* s, double* dst[3], double* const src[3], bool to_dev, int device_ptrs) {
    const long n[3] = {s->nu(), s->nv(), s->nw()};
    const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice
                                           : (to_dev ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost);
    for (int a = 0; a < 3; ++a) HIPCHK(hipMemcpyAsync(dst[a], src[a], n[a] * 8, kind, s->st));
    return 0;
}
int xf_sim_project_fields(xf_sim* s, double* u, double* v, double* w, int device_ptrs) {
    if (check(s) || !u || !v || !w) return set_err(-1, "null argument");
    double* dev[3] = {s->u, s->v, s->w};
    double* const io[3] = {u, v, w};
    int rc = copy_faces(s, dev, io, true, device_ptrs);
    if (rc < 0) return rc;
    int status = st_solve(s);
    if (status < 0) return status;
    if ((rc = st_apply(s)) < 0) return rc;
    double* out[3] = {u, v, w};
    double* const from[3] = {s->u, s->v, s->w};
    if ((rc = copy_faces(s, out, from, false, device_ptrs)) < 0) return rc;
    HIPCHK(hipStreamSynchronize(s->st));
    return status;
}
int xf_sim_project_adjoint(xf_sim* s, double* gu, double* gv, double* gw, int device_ptrs) {
    if (check(s) || !gu || !gv || !gw) return set_err(-1, "null argument");
    if (s->prm.pressure_solver != XF_PRESSURE_VARIATIONAL)
        return set_err(-1, "the projection adjoint is built for the variational pressure operator");
    if (!s->ga_u) {
        if (dalloc(&s->ga_u, s->nu()) != hipSuccess || dalloc(&s->ga_v, s->nv()) != hipSuccess ||
            dalloc(&s->ga_w, s->nw()) != hipSuccess || dalloc(&s->ga_p, s->g.N) != hipSuccess)
            return set_err(-2, "adjoint allocation failed");
    }
    double* dev[3] = {s->ga_u, s->ga_v, s->ga_w};
    double* const io[3] = {gu, gv, gw};
    int rc = copy_faces(s, dev, io, true, device_ptrs);
    if (rc < 0) return rc;
    // pbar = -G^T g, solve A bbar = pbar, g += B^T bbar
    launch_apply_adjoint(s->g, s->mat, s->ga_u, s->ga_v, s->ga_w, s->ga_p, s->st);
    int status = st_solve_from(s, s->ga_u, s->ga_v, s->ga_w, s->ga_p);
    if (status < 0) return status;
    for (int a = 0; a < 3; ++a) launch_rhs_adjoint(s->g, s->mat, s->pw.x, dev[a], a, s->st);
    HIPCHK(hipGetLastError());
    double* out[3] = {gu, gv, gw};
    double* const from[3] = {s->ga_u, s->ga_v, s->ga_w};
    if ((rc = copy_faces(s, out, from, false, device_ptrs)) < 0) return rc;
    HIPCHK(hipStreamSynchronize(s->st));
    return status;
}

// Caller arrays for the particle adjoints: device pointers are used as they
// are; host arrays are staged through device buffers (copied in when `in`,
// copied back on sync() when `out`).
namespace {
struct Staged {
    xf_sim* s;
    double* host;
    double* dev = nullptr;
    long n;
    bool out, owned = false;
    Staged(xf_sim* s_, const double* h, long n_, bool in, bool out_, int device_ptrs, hipError_t* err)
        : s(s_), host(const_cast<double*>(h)), n(n_), out(out_) {
        if (device_ptrs || n == 0) { dev = host; return; }
        owned = true;
        if ((*err = hipMalloc((void**)&dev, n * 8)) != hipSuccess) { dev = nullptr; return; }
        if (in) *err = hipMemcpyAsync(dev, host, n * 8, hipMemcpyHostToDevice, s->st);
        else *err = hipMemsetAsync(dev, 0, n * 8, s->st);
    }
    hipError_t sync() {
        if (owned && out && dev) return hipMemcpyAsync(host, dev, n * 8, hipMemcpyDeviceToHost, s->st);
        return hipSuccess;
    }
    ~Staged() {
        if (owned && dev) {
            (void)hipStreamSynchronize(s->st);
            (void)hipFree(dev);
        }
    }
};
}  // namespace

#define STAGE_OR_FAIL(v, ...)                                          \
    Staged v(s, __VA_ARGS__, device_ptrs, &herr);                      \
    if (herr != hipSuccess) return set_err(-2, "%s", hipGetErrorString(herr))

int xf_sim_sample_adjoint(xf_sim* s, int64_t n, const double* pos, const double* gvel, double* gpos, double* gu,
                          double* gv, double* gw, int saved, int device_ptrs) {
    if (check(s) || n < 0 || (n > 0 && (!pos || !gvel || !gpos)) || !gu || !gv || !gw)
        return set_err(-1, "null argument");
    hipError_t herr = hipSuccess;
    STAGE_OR_FAIL(P, pos, 3 * n, true, false);
    STAGE_OR_FAIL(GV, gvel, 3 * n, true, false);
    STAGE_OR_FAIL(GP, gpos, 3 * n, true, true);
    STAGE_OR_FAIL(GU, gu, s->nu(), true, true);
    STAGE_OR_FAIL(GVV, gv, s->nv(), true, true);
    STAGE_OR_FAIL(GW, gw, s->nw(), true, true);
    const double* F[3] = {saved ? s->us : s->u, saved ? s->vs : s->v, saved ? s->ws : s->w};
    double* G[3] = {GU.dev, GVV.dev, GW.dev};
    launch_sample_adjoint(s->g, F, G, n, P.dev, GV.dev, GP.dev, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(GP.sync()); HIPCHK(GU.sync()); HIPCHK(GVV.sync()); HIPCHK(GW.sync());
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int xf_sim_advect_adjoint(xf_sim* s, int64_t n, const double* pos, double dt, const double* gnew, double* gpos,
                          double* gu, double* gv, double* gw, int device_ptrs) {
    if (check(s) || n < 0 || (n > 0 && (!pos || !gnew || !gpos)) || !gu || !gv || !gw)
        return set_err(-1, "null argument");
    hipError_t herr = hipSuccess;
    STAGE_OR_FAIL(P, pos, 3 * n, true, false);
    STAGE_OR_FAIL(GN, gnew, 3 * n, true, false);
    STAGE_OR_FAIL(GP, gpos, 3 * n, false, true);
    STAGE_OR_FAIL(GU, gu, s->nu(), true, true);
    STAGE_OR_FAIL(GVV, gv, s->nv(), true, true);
    STAGE_OR_FAIL(GW, gw, s->nw(), true, true);
    const double* F[3] = {s->u, s->v, s->w};
    double* G[3] = {GU.dev, GVV.dev, GW.dev};
    launch_advect_adjoint(s->g, F, G, s->mat, n, P.dev, dt, s->prm.integrator, GN.dev, GP.dev, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(GP.sync()); HIPCHK(GU.sync()); HIPCHK(GVV.sync()); HIPCHK(GW.sync());
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int xf_sim_p2g_adjoint(xf_sim* s, int64_t n, const double* pos, const double* vel, const double* gu,
                       const double* gv, const double* gw, double* gvel, double* gpos, int device_ptrs) {
    if (check(s) || n < 0 || (n > 0 && (!pos || !vel || !gvel || !gpos)) || !gu || !gv || !gw)
        return set_err(-1, "null argument");
    hipError_t herr = hipSuccess;
    STAGE_OR_FAIL(P, pos, 3 * n, true, false);
    STAGE_OR_FAIL(V, vel, 3 * n, true, false);
    STAGE_OR_FAIL(GVEL, gvel, 3 * n, false, true);
    STAGE_OR_FAIL(GP, gpos, 3 * n, true, true);
    STAGE_OR_FAIL(GU, gu, s->nu(), true, false);
    STAGE_OR_FAIL(GVV, gv, s->nv(), true, false);
    STAGE_OR_FAIL(GW, gw, s->nw(), true, false);
    const long mf = s->maxface();
    double *wt = nullptr, *val = nullptr;
    if (dalloc(&wt, mf) != hipSuccess || dalloc(&val, mf) != hipSuccess) {
        dfree(wt); dfree(val);
        return set_err(-2, "p2g adjoint scratch");
    }
    const double* gf[3] = {GU.dev, GVV.dev, GW.dev};
    launch_p2g_adjoint(s->g, s->prm.p2g_kernel, n, P.dev, V.dev, gf, GVEL.dev, GP.dev, wt, val, s->st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = GVEL.sync();
    if (e == hipSuccess) e = GP.sync();
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    dfree(wt); dfree(val);
    if (e != hipSuccess) return set_err(-2, "p2g adjoint: %s", hipGetErrorString(e));
    return 0;
}

// The same VJP with no allocation: device pointers only, the weight and value
// sums in the per-GPU static scratch (cleared per face array by a kernel).
int xf_sim_p2g_adjoint_noalloc(xf_sim* s, int64_t n, const double* pos, const double* vel, const double* gu,
                               const double* gv, const double* gw, double* gvel, double* gpos) {
    if (check(s) || n < 0 || (n > 0 && (!pos || !vel || !gvel || !gpos)) || !gu || !gv || !gw)
        return set_err(-1, "null argument");
    const long mf = s->maxface();
    if (mf > XF_P2G_ADJOINT_NOALLOC_MAX_FACES)
        return set_err(-1, "p2g adjoint (no allocation): the grid has %ld faces per axis, the static scratch holds %d",
                       mf, XF_P2G_ADJOINT_NOALLOC_MAX_FACES);
    double* base = scratch_ptr(s->dev);
    if (!base) return set_err(-2, "p2g adjoint (no allocation): no scratch address for device %d", s->dev);
    const double* gf[3] = {gu, gv, gw};
    launch_p2g_adjoint(s->g, s->prm.p2g_kernel, n, pos, vel, gf, gvel, gpos, base,
                       base + XF_P2G_ADJOINT_NOALLOC_MAX_FACES, s->st);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) return set_err(-2, "p2g adjoint (no allocation): %s", hipGetErrorString(e));
    return 0;
}

int xf_sim_extrapolate_adjoint(xf_sim* s, double* gu, double* gv, double* gw, int device_ptrs) {
    if (check(s) || !gu || !gv || !gw) return set_err(-1, "null argument");
    const int layers = (int)(s->prm.cfl + 5);
    if (layers > 120) return set_err(-1, "extrapolation layers %d exceed 120", layers);
    hipError_t herr = hipSuccess;
    STAGE_OR_FAIL(GU, gu, s->nu(), true, true);
    STAGE_OR_FAIL(GVV, gv, s->nv(), true, true);
    STAGE_OR_FAIL(GW, gw, s->nw(), true, true);
    double* G[3] = {GU.dev, GVV.dev, GW.dev};
    for (int a = 0; a < 3; ++a) launch_extrapolate_adjoint(s->g, s->mat, G[a], s->marker, a, layers, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(GU.sync()); HIPCHK(GVV.sync()); HIPCHK(GW.sync());
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}

int xf_sim_density_adjoint(xf_sim* s, int64_t n, const double* pos, const double* gphi, double* gpos,
                           int device_ptrs) {
    if (check(s) || n < 0 || (n > 0 && (!pos || !gpos)) || !gphi) return set_err(-1, "null argument");
    hipError_t herr = hipSuccess;
    const int d = s->prm.sub_division;
    const long nphi = (long)(d * s->g.ni + 1) * (d * s->g.nj + 1) * (d * s->g.nk + 1);
    STAGE_OR_FAIL(P, pos, 3 * n, true, false);
    STAGE_OR_FAIL(GPHI, gphi, nphi, true, false);
    STAGE_OR_FAIL(GP, gpos, 3 * n, true, true);
    launch_density_adjoint(s->g, d, blob_radius(s), n, P.dev, GPHI.dev, GP.dev, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(GP.sync());
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}

int xf_sim_apply_matrix(xf_sim* s, const double* x, double* y) {
    if (check(s) || !x || !y) return set_err(-1, "null argument");
    const long N = s->g.N;
    HIPCHK(hipMemcpyAsync(s->pw.s, x, N * 8, hipMemcpyHostToDevice, s->st));
    pcg_spmv(s->g, s->code, s->pw.s, s->pw.z, s->st);
    HIPCHK(hipMemcpyAsync(y, s->pw.z, N * 8, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}
int xf_sim_apply_precond(xf_sim* s, const double* r, double* z) {
    if (check(s) || !r || !z) return set_err(-1, "null argument");
    const long N = s->g.N;
    HIPCHK(hipMemcpyAsync(s->pw.s, r, N * 8, hipMemcpyHostToDevice, s->st));
    const int sweep = s->prm.sweep_kernel == 0 ? 9 : s->prm.sweep_kernel;  // as in the solve
    pcg_precond(s->g, s->code, s->pw.inv, s->pw.s, s->pw.q, s->pw.z, sweep, s->pw, s->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(z, s->pw.z, N * 8, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return 0;
}

int xf_project(int ni, int nj, int nk, const uint8_t* material, double* u, double* v, double* w, double* pressure,
               const xf_params* p, xf_solve_stats* st) {
    if (!material || !u || !v || !w) return set_err(-1, "null argument");
    xf_sim* s = nullptr;
    int rc = xf_sim_create(ni, nj, nk, &s);
    if (rc) return rc;
    xf_params prm;
    if (p) prm = *p; else xf_params_default(&prm);
    rc = xf_sim_set_params(s, &prm);
    const long N = (long)ni * nj * nk;
    // the reference's variational assembly has no range checks
    // (variational_pressuresolver.cpp:252-292): fluid must not touch the grid boundary
    for (int k = 0; k < nk && rc == 0 && prm.pressure_solver == XF_PRESSURE_VARIATIONAL; ++k)
        for (int j = 0; j < nj && rc == 0; ++j)
            for (int i = 0; i < ni; ++i)
                if ((i == 0 || j == 0 || k == 0 || i == ni - 1 || j == nj - 1 || k == nk - 1) &&
                    material[i + (long)ni * (j + (long)nj * k)] == XF_FLUID) {
                    rc = set_err(-1, "fluid cell on the grid boundary at (%d,%d,%d)", i, j, k);
                    break;
                }
    if (!rc) rc = xf_sim_set_field(s, XF_FIELD_MATERIAL, material, N);
    if (!rc) rc = xf_sim_set_field(s, XF_FIELD_U, u, s->nu() * 8);
    if (!rc) rc = xf_sim_set_field(s, XF_FIELD_V, v, s->nv() * 8);
    if (!rc) rc = xf_sim_set_field(s, XF_FIELD_W, w, s->nw() * 8);
    int status = 0;
    if (!rc) {
        status = xf_sim_stage(s, XF_STAGE_SOLVE, 0.0);
        if (status < 0) rc = status;
    }
    if (!rc) rc = xf_sim_stage(s, XF_STAGE_APPLY, 0.0);
    if (!rc) rc = xf_sim_get_field(s, XF_FIELD_U, u, s->nu() * 8);
    if (!rc) rc = xf_sim_get_field(s, XF_FIELD_V, v, s->nv() * 8);
    if (!rc) rc = xf_sim_get_field(s, XF_FIELD_W, w, s->nw() * 8);
    if (!rc && pressure) rc = xf_sim_get_field(s, XF_FIELD_PRESSURE, pressure, N * 8);
    if (st) *st = s->last;
    xf_sim_destroy(s);
    return rc ? rc : status;
}

}  // extern "C"
