// xf_mesh.hip -- marching-cubes surface mesh of the particle-blob density
// field, in the reference's vertex and triangle order, plus its umbrella
// smoothing and vertex normals.
//
// FluidMesher::generateMesh (fluidmesher.cpp:364-392) walks the sub-surface
// cells in list order; in each cell it visits the cut edges 0..11 and creates a
// vertex for an edge the first time any cell meets it (edge keyed by its lower
// endpoint, fluidmesher.cpp:464-476), then appends the case's triangles.  The
// serial order is reproduced without a serial loop:
//   1. every (cell q, cut edge e) claims its edge with atomicMin(owner, q);
//   2. (q, e) creates the vertex iff owner == q; a prefix sum over the
//      (q-major, e-minor) slots gives each new vertex its serial index;
//   3. owners are replaced by vertex indices, triangles are written at the
//      prefix sum of the per-case triangle counts;
//   4. the touched edge slots are released (the owner array stays allocated).
// Vertex positions use vertexInterp's arithmetic (fluidmesher.cpp:28-37) on
// idx2Pos points ((v + 0.5) * dx / d, xfluidutil.cpp:18-21).
//
// Smoothing (fluidmesher.cpp:547-556 -> trianglemesh.cpp:736-810): 4 Jacobi
// steps v += 0.8 (avg - v), avg over the other corners of the vertex's
// triangles taken in ascending triangle order (a per-vertex list, sorted), so
// the sums match the reference bit for bit.  Normals: trianglemesh.cpp:332-360.
#include <climits>
#include <vector>
#include "xf_internal.h"
#include "xf_mc_tables.h"

__constant__ McTables c_mc = mc_make_tables();

struct MeshGeo {
    long sni, snj;    // sub-cells along i, j
    long VI, VJ, NV;  // vertex lattice of the density field
    double sdx, iso;
};

__device__ inline void sub_cell(const MeshGeo& m, int32_t cell, long* a, long* b, long* c) {
    *a = cell % m.sni;
    *b = (cell / m.sni) % m.snj;
    *c = cell / (m.sni * m.snj);
}
__device__ inline long edge_key(const MeshGeo& m, long a, long b, long c, int e) {
    const int l = kMcEdge[e][0];
    return kMcEdgeDir[e] * m.NV + (a + kMcCorner[l][0]) + m.VI * ((b + kMcCorner[l][1]) + m.VJ * (c + kMcCorner[l][2]));
}

__global__ void k_mesh_claim(MeshGeo m, long slots, const int32_t* __restrict__ cells, const uint8_t* __restrict__ cube,
                             int32_t* owner) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= slots) return;
    const long q = t / 12;
    const int e = (int)(t % 12);
    if (!((c_mc.cut[cube[q]] >> e) & 1)) return;
    long a, b, c;
    sub_cell(m, cells[q], &a, &b, &c);
    atomicMin(&owner[edge_key(m, a, b, c, e)], (int32_t)q);
}

__global__ void k_mesh_flag(MeshGeo m, long slots, const int32_t* __restrict__ cells, const uint8_t* __restrict__ cube,
                            const int32_t* __restrict__ owner, int32_t* flag) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= slots) return;
    const long q = t / 12;
    const int e = (int)(t % 12);
    int f = 0;
    if ((c_mc.cut[cube[q]] >> e) & 1) {
        long a, b, c;
        sub_cell(m, cells[q], &a, &b, &c);
        f = owner[edge_key(m, a, b, c, e)] == (int32_t)q;
    }
    flag[t] = f;
}

// new vertices: position by vertexInterp, owner slot -> vertex index
__global__ void k_mesh_vertices(MeshGeo m, long slots, const int32_t* __restrict__ cells,
                                const uint8_t* __restrict__ cube, const int32_t* __restrict__ flag,
                                const int32_t* __restrict__ vpos, const double* __restrict__ phi, int32_t* owner,
                                double* __restrict__ verts) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= slots || !flag[t]) return;
    const long q = t / 12;
    const int e = (int)(t % 12);
    long a, b, c;
    sub_cell(m, cells[q], &a, &b, &c);
    const int* l = kMcCorner[kMcEdge[e][0]];
    const int* r = kMcCorner[kMcEdge[e][1]];
    const long li = a + l[0], lj = b + l[1], lk = c + l[2];
    const long ri = a + r[0], rj = b + r[1], rk = c + r[2];
    const double v1 = phi[li + m.VI * (lj + m.VJ * lk)], v2 = phi[ri + m.VI * (rj + m.VJ * rk)];
    const double p1[3] = {(li + 0.5) * m.sdx, (lj + 0.5) * m.sdx, (lk + 0.5) * m.sdx};
    const double p2[3] = {(ri + 0.5) * m.sdx, (rj + 0.5) * m.sdx, (rk + 0.5) * m.sdx};
    double s = v1 - v2;
    if (-1e-7 < s && 1e-7 > s) s = 0.5;
    else s = (v1 - m.iso) / s;
    const long v = vpos[t];
    for (int x = 0; x < 3; ++x) verts[3 * v + x] = p1[x] + (p2[x] - p1[x]) * s;
    owner[edge_key(m, a, b, c, e)] = (int32_t)v;
}

__global__ void k_mesh_tri_count(long ncells, const uint8_t* __restrict__ cube, int32_t* cnt) {
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q < ncells) cnt[q] = c_mc.ntri[cube[q]];
}

__global__ void k_mesh_triangles(MeshGeo m, long ncells, const int32_t* __restrict__ cells,
                                 const uint8_t* __restrict__ cube, const int32_t* __restrict__ tpos,
                                 const int32_t* __restrict__ owner, int32_t* __restrict__ tris) {
    const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (q >= ncells) return;
    const int ci = cube[q];
    const int n = 3 * c_mc.ntri[ci];
    if (n == 0) return;
    long a, b, c;
    sub_cell(m, cells[q], &a, &b, &c);
    int32_t* out = tris + 3 * (long)tpos[q];
    for (int x = 0; x < n; ++x) out[x] = owner[edge_key(m, a, b, c, c_mc.tri[ci][x])];
}

__global__ void k_mesh_release(MeshGeo m, long slots, const int32_t* __restrict__ cells,
                               const uint8_t* __restrict__ cube, int32_t* owner) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= slots) return;
    const long q = t / 12;
    const int e = (int)(t % 12);
    if (!((c_mc.cut[cube[q]] >> e) & 1)) return;
    long a, b, c;
    sub_cell(m, cells[q], &a, &b, &c);
    owner[edge_key(m, a, b, c, e)] = INT_MAX;
}

__global__ void k_fill_i32(long n, int32_t* p, int32_t v) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

// ---- vertex -> triangles lists (trianglemesh.cpp:609-630) ----
__global__ void k_vt_count(long ntris, const int32_t* __restrict__ tris, int32_t* cnt) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    for (int x = 0; x < 3; ++x) atomicAdd(&cnt[tris[3 * t + x]], 1);
}
__global__ void k_vt_fill(long ntris, const int32_t* __restrict__ tris, const int32_t* __restrict__ start,
                          int32_t* cursor, int32_t* list) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    for (int x = 0; x < 3; ++x) {
        const int32_t v = tris[3 * t + x];
        list[start[v] + atomicAdd(&cursor[v], 1)] = (int32_t)t;
    }
}
__global__ void k_vt_sort(long nverts, const int32_t* __restrict__ start, int32_t* list) {
    const long v = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (v >= nverts) return;
    const int32_t b = start[v], e = start[v + 1];
    for (int32_t i = b + 1; i < e; ++i) {
        const int32_t x = list[i];
        int32_t j = i - 1;
        while (j >= b && list[j] > x) { list[j + 1] = list[j]; --j; }
        list[j + 1] = x;
    }
}

// TriangleMesh::_smoothTriangleMesh (trianglemesh.cpp:736-775)
__global__ void k_smooth(long nverts, const int32_t* __restrict__ start, const int32_t* __restrict__ list,
                         const int32_t* __restrict__ tris, const double* __restrict__ vin, double* __restrict__ vout,
                         double value) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nverts) return;
    double ax = 0.0, ay = 0.0, az = 0.0;
    int count = 0;
    for (int32_t p = start[i]; p < start[i + 1]; ++p) {
        const int32_t* t = tris + 3 * (long)list[p];
        for (int x = 0; x < 3; ++x) {
            const int32_t o = t[x];
            if (o == (int32_t)i) continue;
            ax = ax + vin[3 * (long)o];
            ay = ay + vin[3 * (long)o + 1];
            az = az + vin[3 * (long)o + 2];
            ++count;
        }
    }
    const double cnt = (double)count;
    const double mx = ax / cnt, my = ay / cnt, mz = az / cnt;
    const double x0 = vin[3 * i], y0 = vin[3 * i + 1], z0 = vin[3 * i + 2];
    vout[3 * i] = x0 + (mx - x0) * value;
    vout[3 * i + 1] = y0 + (my - y0) * value;
    vout[3 * i + 2] = z0 + (mz - z0) * value;
}

// TriangleMesh::updateVertexNormals (trianglemesh.cpp:332-360)
__device__ inline void normalize3(double* x) {
    const double l = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    x[0] = x[0] / l; x[1] = x[1] / l; x[2] = x[2] / l;
}
__global__ void k_face_normals(long ntris, const int32_t* __restrict__ tris, const double* __restrict__ verts,
                               double* __restrict__ fn) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    const double* p0 = verts + 3 * (long)tris[3 * t];
    const double* p1 = verts + 3 * (long)tris[3 * t + 1];
    const double* p2 = verts + 3 * (long)tris[3 * t + 2];
    const double a0 = p1[0] - p0[0], a1 = p1[1] - p0[1], a2 = p1[2] - p0[2];
    const double b0 = p2[0] - p0[0], b1 = p2[1] - p0[1], b2 = p2[2] - p0[2];
    double c[3] = {a1 * b2 - a2 * b1, a2 * b0 - a0 * b2, a0 * b1 - a1 * b0};
    normalize3(c);
    for (int x = 0; x < 3; ++x) fn[3 * t + x] = c[x];
}
__global__ void k_vertex_normals(long nverts, const int32_t* __restrict__ start, const int32_t* __restrict__ list,
                                 const double* __restrict__ fn, double* __restrict__ vn) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nverts) return;
    // the sum starts from zero (the reference's default-constructed Vec3d is
    // uninitialised, trianglemesh.cpp:351)
    double n[3] = {0.0, 0.0, 0.0};
    for (int32_t p = start[i]; p < start[i + 1]; ++p)
        for (int x = 0; x < 3; ++x) n[x] = n[x] + fn[3 * (long)list[p] + x];
    const double cnt = (double)(float)(start[i + 1] - start[i]);
    for (int x = 0; x < 3; ++x) n[x] = n[x] / cnt;
    normalize3(n);
    for (int x = 0; x < 3; ++x) vn[3 * i + x] = n[x];
}

// ---------------------------------------------------------------------------
namespace {
template <class T>
hipError_t grow(T** p, long* cap, long need) {
    if (*p && *cap >= need) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    long n = need < 1024 ? 1024 : need + need / 4;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    *cap = e == hipSuccess ? n : 0;
    return e;
}
int32_t read_i32(const int32_t* d, hipStream_t st) {
    int32_t v = 0;
    if (hipMemcpyAsync(&v, d, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return -1;
    return v;
}
}  // namespace

void mesh_free(MeshWork& w) {
    void* ps[] = {w.owner, w.cells, w.cube, w.flag, w.vpos, w.tcnt, w.tpos, w.verts, w.verts2, w.normals, w.fnorm,
                  w.tris, w.vt_cnt, w.vt_start, w.vt_cursor, w.vt_list, w.scan_tmp};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    w = MeshWork{};
}

// Mesh of phi (density field of sub-division d) -> w.verts / w.tris / w.normals.
int build_mesh(const Geo& g, const double* phi, int d, double iso, int smooth_iters, double smooth_value,
               int32_t* count_tmp, int32_t* start_tmp, int32_t* scan_tmp, int scan_tmp_n, MeshWork& w,
               hipStream_t st) {
    MeshGeo m;
    m.sni = (long)d * g.ni;
    m.snj = (long)d * g.nj;
    m.VI = m.sni + 1;
    m.VJ = m.snj + 1;
    m.NV = m.VI * m.VJ * ((long)d * g.nk + 1);
    m.sdx = g.dx / (double)d;
    m.iso = iso;
    if (3 * m.NV > INT_MAX) return -1;
    // owner slots of all edges, free = INT_MAX (released after each mesh)
    if (!w.owner || w.nowner != 3 * m.NV) {
        if (w.owner) (void)hipFree(w.owner);
        w.owner = nullptr;
        if (hipMalloc((void**)&w.owner, 3 * m.NV * sizeof(int32_t)) != hipSuccess) return -2;
        w.nowner = 3 * m.NV;
        k_fill_i32<<<1024, XF_BLOCK, 0, st>>>(w.nowner, w.owner, INT_MAX);
    }
    // sub-surface cells in the reference's order
    const long S = surface_cells(g, phi, d, iso, nullptr, nullptr, 0, count_tmp, start_tmp, scan_tmp, scan_tmp_n, st);
    w.ncells = S;
    w.nverts = 0;
    w.ntris = 0;
    if (S == 0) return 0;
    if (grow(&w.cells, &w.cap_cells, S) != hipSuccess || grow(&w.cube, &w.cap_cube, S) != hipSuccess) return -2;
    surface_cells(g, phi, d, iso, w.cells, w.cube, S, count_tmp, start_tmp, scan_tmp, scan_tmp_n, st);
    const long slots = 12 * S;
    if (grow(&w.flag, &w.cap_flag, slots) != hipSuccess || grow(&w.vpos, &w.cap_vpos, slots + 1) != hipSuccess ||
        grow(&w.tcnt, &w.cap_tcnt, S) != hipSuccess || grow(&w.tpos, &w.cap_tpos, S + 1) != hipSuccess ||
        grow(&w.scan_tmp, &w.cap_scan, scan_tmp_size(slots + 1)) != hipSuccess)
        return -2;
    k_mesh_claim<<<xf_blocks(slots), XF_BLOCK, 0, st>>>(m, slots, w.cells, w.cube, w.owner);
    k_mesh_flag<<<xf_blocks(slots), XF_BLOCK, 0, st>>>(m, slots, w.cells, w.cube, w.owner, w.flag);
    exclusive_scan_i32(w.flag, w.vpos, slots, w.scan_tmp, (int)w.cap_scan, st);
    k_mesh_tri_count<<<xf_blocks(S), XF_BLOCK, 0, st>>>(S, w.cube, w.tcnt);
    exclusive_scan_i32(w.tcnt, w.tpos, S, w.scan_tmp, (int)w.cap_scan, st);
    const long nv = read_i32(w.vpos + slots, st);
    const long nt = read_i32(w.tpos + S, st);
    if (nv < 0 || nt < 0) return -2;
    if (grow(&w.verts, &w.cap_verts, 3 * nv + 3) != hipSuccess ||
        grow(&w.verts2, &w.cap_verts2, 3 * nv + 3) != hipSuccess ||
        grow(&w.normals, &w.cap_normals, 3 * nv + 3) != hipSuccess ||
        grow(&w.tris, &w.cap_tris, 3 * nt + 3) != hipSuccess || grow(&w.fnorm, &w.cap_fnorm, 3 * nt + 3) != hipSuccess)
        return -2;
    k_mesh_vertices<<<xf_blocks(slots), XF_BLOCK, 0, st>>>(m, slots, w.cells, w.cube, w.flag, w.vpos, phi, w.owner,
                                                           w.verts);
    k_mesh_triangles<<<xf_blocks(S), XF_BLOCK, 0, st>>>(m, S, w.cells, w.cube, w.tpos, w.owner, w.tris);
    k_mesh_release<<<xf_blocks(slots), XF_BLOCK, 0, st>>>(m, slots, w.cells, w.cube, w.owner);
    w.nverts = nv;
    w.ntris = nt;
    if (nv == 0) return 0;
    // vertex -> triangle lists in ascending triangle order
    if (grow(&w.vt_cnt, &w.cap_vt_cnt, nv) != hipSuccess || grow(&w.vt_start, &w.cap_vt_start, nv + 1) != hipSuccess ||
        grow(&w.vt_cursor, &w.cap_vt_cursor, nv) != hipSuccess || grow(&w.vt_list, &w.cap_vt_list, 3 * nt) != hipSuccess ||
        grow(&w.scan_tmp, &w.cap_scan, scan_tmp_size(nv + 1)) != hipSuccess)
        return -2;
    k_fill_i32<<<xf_blocks(nv), XF_BLOCK, 0, st>>>(nv, w.vt_cnt, 0);
    k_fill_i32<<<xf_blocks(nv), XF_BLOCK, 0, st>>>(nv, w.vt_cursor, 0);
    k_vt_count<<<xf_blocks(nt), XF_BLOCK, 0, st>>>(nt, w.tris, w.vt_cnt);
    exclusive_scan_i32(w.vt_cnt, w.vt_start, nv, w.scan_tmp, (int)w.cap_scan, st);
    k_vt_fill<<<xf_blocks(nt), XF_BLOCK, 0, st>>>(nt, w.tris, w.vt_start, w.vt_cursor, w.vt_list);
    k_vt_sort<<<xf_blocks(nv), XF_BLOCK, 0, st>>>(nv, w.vt_start, w.vt_list);
    for (int it = 0; it < smooth_iters; ++it) {
        k_smooth<<<xf_blocks(nv), XF_BLOCK, 0, st>>>(nv, w.vt_start, w.vt_list, w.tris, w.verts, w.verts2,
                                                     smooth_value);
        std::swap(w.verts, w.verts2);
        std::swap(w.cap_verts, w.cap_verts2);
    }
    k_face_normals<<<xf_blocks(nt), XF_BLOCK, 0, st>>>(nt, w.tris, w.verts, w.fnorm);
    k_vertex_normals<<<xf_blocks(nv), XF_BLOCK, 0, st>>>(nv, w.vt_start, w.vt_list, w.fnorm, w.normals);
    return 0;
}
