// Triangle-mesh particle boundary: the reference's triangle_mesh_trigger
// (particle_boundary/triangle_mesh_trigger.hpp:30-437) and boundary_router
// with basic_material_set (boundary_router.hpp:20-92,
// basic_material_set.hpp:15-119).
//
// The partition is built once on the host: triangles get unit normals
// (zero-area ones are dropped), and each is listed in every cell of a
// uniform I x J x K partition of the mesh's box whose cell box it overlaps,
// in triangle order.  On the device it is CSR: cell_start[I*J*K + 1],
// cell_tri[] (triangle indices) and the triangles as 12 doubles (vertices,
// normal) with an int32 id each.  The segment kernels run one thread per
// particle: the segment old -> new position is traversed cell by cell from
// its start, and the first cell with a hit returns the smallest segment
// parameter among its triangles (the reference's restrict mode); the fast
// mode tests only the start and end cells.  The router applies the hit
// triangle's material (its id): diffuse (cosine-law direction), specular,
// Maxwell re-emission (half-Maxwellian) or absorption (position x = NaN).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <thread>
#include <vector>

#include "common.hpp"
#include "psum_mesh_geom.hpp"

namespace psum_hip {
namespace {

constexpr int kMeshThreads = 256;
inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

struct Mesh {
    int I = 0, J = 0, K = 0;
    int fast = 0;
    double box[6] = {0, 0, 0, 0, 0, 0};
    double rd[3] = {0, 0, 0};
    int64_t kept = 0, entries = 0, ntri = 0;
    int32_t max_id = -1;
    std::vector<uint32_t> h_start, h_tri;
    std::vector<double> h_tri12;
    uint32_t* cell_start = nullptr;
    uint32_t* cell_tri = nullptr;
    double* tri12 = nullptr;
    int32_t* ids = nullptr;
    std::vector<float> h_udf;
    float* udf = nullptr;
    int udf_on = 0;

    ~Mesh()
    {
        if (udf) (void)hipFree(udf);
        if (cell_start) (void)hipFree(cell_start);
        if (cell_tri) (void)hipFree(cell_tri);
        if (tri12) (void)hipFree(tri12);
        if (ids) (void)hipFree(ids);
    }

    psum_mesh_view view() const
    {
        psum_mesh_view v{};
        v.I = I;
        v.J = J;
        v.K = K;
        v.fast = fast;
        for (int c = 0; c < 6; ++c) v.box[c] = box[c];
        for (int c = 0; c < 3; ++c) v.rd[c] = rd[c];
        v.cell_start = cell_start;
        v.cell_tri = cell_tri;
        v.tri12 = tri12;
        v.ids = ids;
        v.udf = udf_on ? udf : nullptr;
        return v;
    }
};

// the automatic box: the triangles' bounding box widened by the
// reference's fractions (triangle_mesh_trigger.hpp:378-392)
void auto_box(const double* tri, int64_t ntri, double* box)
{
    double lo[3] = {tri[0], tri[1], tri[2]}, hi[3] = {tri[0], tri[1], tri[2]};
    for (int64_t i = 0; i < ntri; ++i)
        for (int v = 0; v < 3; ++v)
            for (int c = 0; c < 3; ++c) {
                const double x = tri[9 * i + 3 * v + c];
                lo[c] = std::min(lo[c], x);
                hi[c] = std::max(hi[c], x);
            }
    const double fl[3] = {0.1003, 0.1027, 0.1047}, fh[3] = {0.1057, 0.1059, 0.1069};
    for (int c = 0; c < 3; ++c) {
        box[c] = lo[c] - fl[c] * (hi[c] - lo[c]);
        box[3 + c] = hi[c] + fh[c] * (hi[c] - lo[c]);
    }
}

// triangle_mesh_trigger::set (:171-275): normals, bounding-box cell range,
// box/triangle overlap per cell; per-cell lists in triangle order
void build(Mesh& m, const double* tri, const int32_t* ids, int64_t ntri)
{
    const int64_t ncell = (int64_t)m.I * m.J * m.K;
    const double span[3] = {m.box[3] - m.box[0], m.box[4] - m.box[1], m.box[5] - m.box[2]};
    for (int c = 0; c < 3; ++c)
        if (!(span[c] >= mesh::kSmall)) throw std::invalid_argument("psum_mesh_create: box has a zero or negative span");
    const int dims[3] = {m.I, m.J, m.K};
    for (int c = 0; c < 3; ++c) m.rd[c] = (double)dims[c] / span[c];
    m.h_tri12.assign((size_t)ntri * 12, 0.0);
    std::vector<std::vector<uint32_t>> lists((size_t)ncell);
    const double dx = span[0] / (double)m.I, dy = span[1] / (double)m.J, dz = span[2] / (double)m.K;
    for (int64_t i = 0; i < ntri; ++i) {
        const double* t = tri + 9 * i;
        const double e1x = t[3] - t[0], e1y = t[4] - t[1], e1z = t[5] - t[2];
        const double e2x = t[6] - t[0], e2y = t[7] - t[1], e2z = t[8] - t[2];
        const double nx = e1y * e2z - e1z * e2y;
        const double ny = e1z * e2x - e1x * e2z;
        const double nz = e1x * e2y - e1y * e2x;
        const double nrm = std::sqrt(nx * nx + ny * ny + nz * nz);
        if (nrm < mesh::kSmall) continue;     // zero-area triangle: dropped
        ++m.kept;
        double* o = &m.h_tri12[(size_t)i * 12];
        std::memcpy(o, t, 9 * sizeof(double));
        o[9] = nx / nrm;
        o[10] = ny / nrm;
        o[11] = nz / nrm;
        m.max_id = std::max(m.max_id, ids[i]);
        int lo[3], hi[3];
        for (int c = 0; c < 3; ++c) {
            const double mn = std::fmin(t[c], std::fmin(t[3 + c], t[6 + c]));
            const double mx = std::fmax(t[c], std::fmax(t[3 + c], t[6 + c]));
            const double a = std::floor((mn - m.box[c]) * m.rd[c]), b = std::ceil((mx - m.box[c]) * m.rd[c]);
            lo[c] = a < 0 ? 0 : (a >= dims[c] ? dims[c] - 1 : (int)a);
            hi[c] = b < 0 ? 0 : (b >= dims[c] ? dims[c] - 1 : (int)b);
        }
        for (int a = lo[0]; a <= hi[0]; ++a)
            for (int b = lo[1]; b <= hi[1]; ++b)
                for (int c = lo[2]; c <= hi[2]; ++c) {
                    const double bx0 = m.box[0] + a * dx, by0 = m.box[1] + b * dy, bz0 = m.box[2] + c * dz;
                    const double cb[6] = {bx0, by0, bz0, bx0 + dx, by0 + dy, bz0 + dz};
                    if (mesh::box_tri_overlap(t, cb))
                        lists[(size_t)(((int64_t)a * m.J + b) * m.K + c)].push_back((uint32_t)i);
                }
    }
    m.h_start.assign((size_t)ncell + 1, 0);
    for (int64_t c = 0; c < ncell; ++c) m.h_start[c + 1] = m.h_start[c] + (uint32_t)lists[c].size();
    m.entries = m.h_start[ncell];
    if (m.entries > (int64_t)0xFFFFFFFF) throw std::invalid_argument("psum_mesh_create: partition too large");
    m.h_tri.reserve((size_t)m.entries);
    for (auto& l : lists) m.h_tri.insert(m.h_tri.end(), l.begin(), l.end());
}


// compute_udf (geometry.hpp:343-460): per empty cell the smallest
// cell-distance lower bound to an occupied cell within +-udf_k cells (cells d
// apart share no point closer than max(|d| - 1, 0) per axis), capped at
// udf_k - 1; then passes that raise a cell by one where no neighbour is
// lower, until nothing changes or max_loop + 1 passes
// f(i) for i in [0, n) on up to 16 host threads (each i independent)
template <typename F>
void host_parallel_for(int n, F&& f)
{
    const int nt = std::max(1, std::min({n, 16, (int)std::thread::hardware_concurrency()}));
    psum_hip::host_threads((unsigned)nt, [&](unsigned t) {
        for (int i = (int)t; i < n; i += nt) f(i);
    });
}

void compute_udf(int I, int J, int K, const std::vector<uint32_t>& start, std::vector<float>& udf, int udf_k = 6,
                 int max_loop = 200)
{
    const int64_t n = (int64_t)I * J * K;
    udf.assign((size_t)n, 0.0f);
    std::vector<uint8_t> occ((size_t)n);
    bool any = false;
    for (int64_t c = 0; c < n; ++c) any |= (occ[c] = start[c + 1] > start[c]) != 0;
    if (!any) {
        std::fill(udf.begin(), udf.end(), (float)(std::min(I, std::min(J, K)) - 1));
        return;
    }
    const float cap = (float)(udf_k - 1);
    auto at = [&](int i, int j, int k) { return ((int64_t)i * J + j) * K + k; };
    host_parallel_for(I, [&](int i) {
        for (int j = 0; j < J; ++j)
            for (int k = 0; k < K; ++k) {
                const int64_t c = at(i, j, k);
                if (occ[c]) continue;
                float best = 1e30f;
                for (int a = std::max(0, i - udf_k); a <= std::min(I - 1, i + udf_k); ++a)
                    for (int b = std::max(0, j - udf_k); b <= std::min(J - 1, j + udf_k); ++b)
                        for (int e = std::max(0, k - udf_k); e <= std::min(K - 1, k + udf_k); ++e) {
                            if (!occ[at(a, b, e)]) continue;
                            const int li = std::max(std::abs(a - i) - 1, 0), lj = std::max(std::abs(b - j) - 1, 0),
                                      lk = std::max(std::abs(e - k) - 1, 0);
                            best = std::min(best, std::sqrt((float)(li * li + lj * lj + lk * lk)));
                        }
                udf[c] = std::min(best < 1e29f ? best : cap, cap);
            }
    });
    std::vector<float> nu((size_t)n);
    for (int loop = 0;; ++loop) {
        std::atomic<bool> updated{false};
        host_parallel_for(I, [&](int i) {
            for (int j = 0; j < J; ++j)
                for (int k = 0; k < K; ++k) {
                    const int64_t c = at(i, j, k);
                    const float cur = udf[c];
                    nu[c] = cur;
                    if (!(cur > 0.0f)) continue;
                    bool raise = true;
                    int nb = 0;
                    for (int a = std::max(0, i - 1); a <= std::min(I - 1, i + 1); ++a)
                        for (int b = std::max(0, j - 1); b <= std::min(J - 1, j + 1); ++b)
                            for (int e = std::max(0, k - 1); e <= std::min(K - 1, k + 1); ++e) {
                                if (a == i && b == j && e == k) continue;
                                ++nb;
                                if (udf[at(a, b, e)] < cur) raise = false;
                            }
                    if (raise && nb > 0) {
                        nu[c] = cur + 1.0f;
                        updated = true;
                    }
                }
        });
        udf.swap(nu);
        if (!updated || loop + 1 > max_loop) break;
    }
}

struct SegArgs {
    mesh::View m;
    const void* p1[3];
    const void* p2[3];
    int64_t n;
    double* k;
    double* nrm;
    int32_t* id;
};

template <typename T>
__global__ __launch_bounds__(kMeshThreads) void k_mesh_detect(SegArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * kMeshThreads + threadIdx.x;
    if (i >= a.n) return;
    double p1[3], p2[3];
    bool finite = true;
    for (int c = 0; c < 3; ++c) {
        p1[c] = (double)static_cast<const T*>(a.p1[c])[i];
        p2[c] = (double)static_cast<const T*>(a.p2[c])[i];
        finite = finite && isfinite(p1[c]) && isfinite(p2[c]);
    }
    double k = 0, nrm[3] = {0, 0, 0};
    int32_t id = -1;
    const bool hit = finite && mesh::detect(a.m, p1, p2, k, nrm, id);
    if (a.k) a.k[i] = hit ? k : 0.0;
    if (a.nrm)
        for (int c = 0; c < 3; ++c) a.nrm[3 * i + c] = hit ? nrm[c] : 0.0;
    a.id[i] = hit ? id : -1;
}

struct RouteArgs {
    mesh::View m;
    const void* old[3];
    void* pos[3];
    void* vel[3];
    uint32_t* seed;
    int64_t n;
    double temperature, mass;
    int32_t* hit_id;
    unsigned long long* hits;
};

// boundary_router_acc::deal (boundary_router.hpp:32-58) for every valid
// particle (pos_x_nan_is_invalid)
template <typename T>
__global__ __launch_bounds__(kMeshThreads) void k_mesh_boundary(RouteArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * kMeshThreads + threadIdx.x;
    int hit = 0;
    if (i < a.n) {
        int32_t id = -1;
        double b[3];
        for (int c = 0; c < 3; ++c) b[c] = (double)static_cast<const T*>(a.pos[c])[i];
        if (isfinite(b[0])) {
            double p1[3];
            bool finite = isfinite(b[1]) && isfinite(b[2]);
            for (int c = 0; c < 3; ++c) {
                p1[c] = (double)static_cast<const T*>(a.old[c])[i];
                finite = finite && isfinite(p1[c]);
            }
            double k, nrm[3];
            if (finite && mesh::detect(a.m, p1, b, k, nrm, id)) {
                hit = 1;
                double v[3], h[3];
                for (int c = 0; c < 3; ++c) {
                    v[c] = (double)static_cast<const T*>(a.vel[c])[i];
                    h[c] = p1[c] + k * (b[c] - p1[c]);
                }
                uint32_t s = a.seed[i];
                mesh::material_action(id, b, v, s, h, nrm, a.temperature, a.mass);
                for (int c = 0; c < 3; ++c) {
                    static_cast<T*>(a.pos[c])[i] = (T)b[c];
                    static_cast<T*>(a.vel[c])[i] = (T)v[c];
                }
                a.seed[i] = s;
            }
        }
        if (a.hit_id) a.hit_id[i] = hit ? id : -1;
    }
    if (a.hits) {
        // one atomic per wavefront
        const unsigned long long bal = __ballot(hit);
        if ((threadIdx.x & 63) == 0 && bal) atomicAdd(a.hits, (unsigned long long)__popcll(bal));
    }
}

}  // namespace
}  // namespace psum_hip

using namespace psum_hip;

extern "C" {

void* psum_mesh_create(const double* tri, const int32_t* ids, int64_t ntri, int I, int J, int K, const double* box)
{
    if (!tri || !ids || ntri < 1 || ntri > (int64_t)0x7FFFFFFF) {
        set_error("psum_mesh_create: need 1..2^31-1 triangles and their ids");
        return nullptr;
    }
    if (I < 1 || J < 1 || K < 1 || (int64_t)I * J * K >= (int64_t)0x7FFFFFFF) {
        set_error("psum_mesh_create: I, J, K >= 1 with I*J*K < 2^31-1");
        return nullptr;
    }
    Mesh* m = new Mesh();
    try {
        m->I = I;
        m->J = J;
        m->K = K;
        m->ntri = ntri;
        if (box) std::memcpy(m->box, box, sizeof(m->box));
        else auto_box(tri, ntri, m->box);
        build(*m, tri, ids, ntri);
        const size_t ns = m->h_start.size() * sizeof(uint32_t), nt = std::max<size_t>(m->h_tri.size(), 1) * 4;
        auto chk = [](hipError_t e) {
            if (e != hipSuccess) throw std::runtime_error(std::string("psum_mesh_create: ") + hipGetErrorString(e));
        };
        chk(hipMalloc(&m->cell_start, ns));
        chk(hipMalloc(&m->cell_tri, nt));
        chk(hipMalloc(&m->tri12, (size_t)ntri * 12 * sizeof(double)));
        chk(hipMalloc(&m->ids, (size_t)ntri * sizeof(int32_t)));
        chk(hipMemcpy(m->cell_start, m->h_start.data(), ns, hipMemcpyHostToDevice));
        if (!m->h_tri.empty()) chk(hipMemcpy(m->cell_tri, m->h_tri.data(), m->h_tri.size() * 4, hipMemcpyHostToDevice));
        chk(hipMemcpy(m->tri12, m->h_tri12.data(), (size_t)ntri * 12 * sizeof(double), hipMemcpyHostToDevice));
        chk(hipMemcpy(m->ids, ids, (size_t)ntri * sizeof(int32_t), hipMemcpyHostToDevice));
    } catch (const std::exception& e) {
        set_error(e.what());
        delete m;
        return nullptr;
    }
    return m;
}

psum_status psum_mesh_info(void* h, int64_t* info, double* box)
{
    PSUM_ARG_CHECK(h, "psum_mesh_info: null mesh");
    const Mesh* m = static_cast<const Mesh*>(h);
    if (info) {
        info[0] = m->kept;
        info[1] = m->entries;
        info[2] = (int64_t)m->I * m->J * m->K;
        info[3] = m->max_id;
    }
    if (box) std::memcpy(box, m->box, sizeof(m->box));
    return PSUM_OK;
}

psum_status psum_mesh_partition(void* h, uint32_t* cell_start, uint32_t* cell_tri, double* tri12)
{
    PSUM_ARG_CHECK(h, "psum_mesh_partition: null mesh");
    const Mesh* m = static_cast<const Mesh*>(h);
    if (cell_start) std::memcpy(cell_start, m->h_start.data(), m->h_start.size() * 4);
    if (cell_tri && !m->h_tri.empty()) std::memcpy(cell_tri, m->h_tri.data(), m->h_tri.size() * 4);
    if (tri12) std::memcpy(tri12, m->h_tri12.data(), m->h_tri12.size() * sizeof(double));
    return PSUM_OK;
}

psum_status psum_mesh_set_fast(void* h, int fast)
{
    PSUM_ARG_CHECK(h, "psum_mesh_set_fast: null mesh");
    static_cast<Mesh*>(h)->fast = fast ? 1 : 0;
    return PSUM_OK;
}

psum_status psum_hip_mesh_detect(void* h, int dtype, const void* const* p1, const void* const* p2, int64_t n,
                                 double* k, double* normal, int32_t* id, void* stream)
{
    PSUM_ARG_CHECK(h, "psum_hip_mesh_detect: null mesh");
    PSUM_ARG_CHECK(dtype == PSUM_F64 || dtype == PSUM_F32, "psum_hip_mesh_detect: bad dtype");
    PSUM_ARG_CHECK(n >= 0 && n <= (int64_t)0xFFFFFF00, "psum_hip_mesh_detect: bad particle count");
    if (n == 0) return PSUM_OK;
    PSUM_ARG_CHECK(p1 && p2 && id, "psum_hip_mesh_detect: null pointer");
    SegArgs a{};
    a.m = mesh::view_of(static_cast<const Mesh*>(h)->view());
    for (int c = 0; c < 3; ++c) {
        PSUM_ARG_CHECK(p1[c] && p2[c], "psum_hip_mesh_detect: null coordinate array");
        a.p1[c] = p1[c];
        a.p2[c] = p2[c];
    }
    a.n = n;
    a.k = k;
    a.nrm = normal;
    a.id = id;
    const unsigned nb = (unsigned)((n + kMeshThreads - 1) / kMeshThreads);
    if (dtype == PSUM_F64)
        hipLaunchKernelGGL(k_mesh_detect<double>, dim3(nb), dim3(kMeshThreads), 0, S(stream), a);
    else
        hipLaunchKernelGGL(k_mesh_detect<float>, dim3(nb), dim3(kMeshThreads), 0, S(stream), a);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

psum_status psum_hip_mesh_boundary(void* h, int dtype, const void* const* old, void* const* pos, void* const* vel,
                                   uint32_t* seed, int64_t n, double temperature, double mass, int32_t* hit_id,
                                   unsigned long long* hits, void* stream)
{
    PSUM_ARG_CHECK(h, "psum_hip_mesh_boundary: null mesh");
    PSUM_ARG_CHECK(dtype == PSUM_F64 || dtype == PSUM_F32, "psum_hip_mesh_boundary: bad dtype");
    PSUM_ARG_CHECK(n >= 0 && n <= (int64_t)0xFFFFFF00, "psum_hip_mesh_boundary: bad particle count");
    const Mesh* m = static_cast<const Mesh*>(h);
    PSUM_ARG_CHECK(m->max_id <= PSUM_MAT_ABSORB, "psum_hip_mesh_boundary: triangle ids must be materials 0..3");
    if (n == 0) return PSUM_OK;
    PSUM_ARG_CHECK(old && pos && vel && seed, "psum_hip_mesh_boundary: null pointer");
    RouteArgs a{};
    a.m = mesh::view_of(m->view());
    for (int c = 0; c < 3; ++c) {
        PSUM_ARG_CHECK(old[c] && pos[c] && vel[c], "psum_hip_mesh_boundary: null coordinate array");
        a.old[c] = old[c];
        a.pos[c] = pos[c];
        a.vel[c] = vel[c];
    }
    a.seed = seed;
    a.n = n;
    a.temperature = temperature;
    a.mass = mass;
    a.hit_id = hit_id;
    a.hits = hits;
    const unsigned nb = (unsigned)((n + kMeshThreads - 1) / kMeshThreads);
    if (dtype == PSUM_F64)
        hipLaunchKernelGGL(k_mesh_boundary<double>, dim3(nb), dim3(kMeshThreads), 0, S(stream), a);
    else
        hipLaunchKernelGGL(k_mesh_boundary<float>, dim3(nb), dim3(kMeshThreads), 0, S(stream), a);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

psum_status psum_mesh_get_view(void* h, psum_mesh_view* out)
{
    PSUM_ARG_CHECK(h && out, "psum_mesh_get_view: null pointer");
    *out = static_cast<const Mesh*>(h)->view();
    return PSUM_OK;
}

psum_status psum_mesh_set_udf(void* h, int enabled)
{
    PSUM_ARG_CHECK(h, "psum_mesh_set_udf: null mesh");
    Mesh* m = static_cast<Mesh*>(h);
    if (enabled && !m->udf) {
        try {
            compute_udf(m->I, m->J, m->K, m->h_start, m->h_udf);
        } catch (const std::exception& e) {
            set_error(std::string("psum_mesh_set_udf: ") + e.what());
            return PSUM_ERR_NOMEM;
        }
        PSUM_HIP_TRY(hipMalloc(&m->udf, m->h_udf.size() * sizeof(float)));
        PSUM_HIP_TRY(hipMemcpy(m->udf, m->h_udf.data(), m->h_udf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    m->udf_on = enabled ? 1 : 0;
    return PSUM_OK;
}

void psum_mesh_destroy(void* h) { delete static_cast<Mesh*>(h); }

}  // extern "C"
