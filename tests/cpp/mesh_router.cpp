// The triangle-mesh boundary through the C++ API (psum/psum.hpp:
// mesh_generator, triangle_mesh_trigger, basic_material_set,
// boundary_router), compared on the GPU with the C ABI kernels on the same
// particles:
//   detect    trigger accessor in a kernel (fast, restrict, UDF)  vs  psum_hip_mesh_detect
//   router    drift + boundary_router::deal in for_each          vs  psum_hip_mesh_boundary
// Both sides run the same device code, so every comparison is bit-exact.
// Prints one line per check and "ALL OK" at the end; exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "psum/psum.hpp"

using namespace psum::prelude;
using property::position;
using property::random_seed;
using property::velocity;
using property::weight;
// weight holds the particle's index, so results can be laid out in order
using Particle = tagged_struct<tag_bind<position, vec<3>>, tag_bind<velocity, vec<3>>,
                               tag_bind<random_seed, uint32_t>, tag_bind<weight, double>>;
using Materials = basic_material_set<Particle, pos_x_nan_is_invalid>;
using Mat = Materials::material_type;
using Router = boundary_router<Materials, Particle>;
using Group = particle_group<Particle, pos_x_nan_is_invalid>;

static int fails = 0;
static void report(const char* what, bool ok, double v)
{
    std::printf("%-44s %s (%.6g)\n", what, ok ? "ok" : "FAIL", v);
    if (!ok) ++fails;
}

template <typename T>
static T* dev_copy(const std::vector<T>& h)
{
    T* d = nullptr;
    psum::hip_check(hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(T)), "malloc");
    if (!h.empty()) psum::hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "copy");
    return d;
}

template <typename T>
static std::vector<T> host_copy(const T* d, size_t n)
{
    std::vector<T> h(n);
    if (n) psum::hip_check(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "copy");
    return h;
}

template <typename Acc>
__global__ void k_detect(Acc acc, const double* p1, const double* p2, size_t n, double* k, double* nrm, int32_t* id)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double kk = 0, nx = 0, ny = 0, nz = 0;
    Mat mid = Mat::absorb;
    const bool hit = acc.detect(p1[i], p1[n + i], p1[2 * n + i], p2[i], p2[n + i], p2[2 * n + i], kk, nx, ny, nz, mid);
    k[i] = hit ? kk : 0.0;
    nrm[3 * i] = hit ? nx : 0.0;
    nrm[3 * i + 1] = hit ? ny : 0.0;
    nrm[3 * i + 2] = hit ? nz : 0.0;
    id[i] = hit ? (int32_t)mid : -1;
}

static std::vector<triangle_mesh_with_id<Mat>> scene()
{
    std::vector<std::pair<double, double>> noz = {{-4, 1.0}, {-3, 0.9}, {-2, 0.6}, {-1, 0.6}, {0, 1.0},
                                                  {1, 1.4},  {2, 1.8},  {3, 2.0},  {4, 2.0}};
    return {{mesh_generator::axis_aligned_cube(-5, -5, -5, 5, 5, 10), Mat::absorb},
            {mesh_generator::revolve_zr_curve(noz, 24), Mat::specular},
            {mesh_generator::axis_aligned_plane(-5, -5, -4.0, 5, 5, -4.0), Mat::diffuse},
            {mesh_generator::axis_aligned_plane(-5, 4.0, -5, 5, 4.0, 10), Mat::maxwell_reflect}};
}

static void check_generators(const char* stl)
{
    auto cube = mesh_generator::axis_aligned_cube(0, 0, 0, 1, 2, 3);
    report("generator: cube triangles", cube.size() == 12, (double)cube.size());
    auto rev = mesh_generator::revolve_zr_curve({{0, 1}, {1, 2}, {3, 2}}, 16);
    report("generator: revolve triangles", rev.size() == 2 * 16 * 2, (double)rev.size());
    auto ext = mesh_generator::extrude_xy_line_to_width({{0, 0}, {1, 0}, {1, 1}}, 2.0);
    report("generator: extrude triangles", ext.size() == 8, (double)ext.size());
    auto c2 = mesh_generator::axis_aligned_cube(0, 0, 0, 2, 2, 2);
    mesh_generator::rescale_triangles(c2, 0.5);
    auto [lo, hi] = stl_loader::box_of_triangles(c2);
    report("generator: rescale about the centroid", lo[0] == 0.5 && hi[2] == 1.5, lo[0]);
    auto race = stl_loader::load(stl);
    report("stl: binary fixture triangle count", race.size() == 1952, (double)race.size());
}

int main(int argc, char** argv)
{
    const std::string stl = argc > 1 ? argv[1] : "tests/golden/race.stl";
    check_generators(stl.c_str());
    hipStream_t st;
    psum::hip_check(hipStreamCreate(&st), "stream");
    const auto meshes = scene();
    const int I = 16, J = 16, K = 20;

    // ---- detect: trigger accessor vs psum_hip_mesh_detect, both modes
    {
        triangle_mesh_trigger<Mat> trig(st);
        trig.set(meshes, I, J, K);
        std::mt19937_64 rng(21);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        std::normal_distribution<double> G(0.0, 1.5);
        const size_t n = 20000;
        std::vector<double> p1(3 * n), p2(3 * n);
        const double lo[3] = {-4.9, -4.9, -4.9}, hi[3] = {4.9, 4.9, 9.9};
        for (size_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) {
                p1[c * n + i] = lo[c] + (hi[c] - lo[c]) * U(rng);
                p2[c * n + i] = p1[c * n + i] + G(rng);
            }
        double* d1 = dev_copy(p1);
        double* d2 = dev_copy(p2);
        double *ka, *kb, *na, *nb;
        int32_t *ia, *ib;
        psum::hip_check(hipMalloc(&ka, n * 8), "m");
        psum::hip_check(hipMalloc(&kb, n * 8), "m");
        psum::hip_check(hipMalloc(&na, 3 * n * 8), "m");
        psum::hip_check(hipMalloc(&nb, 3 * n * 8), "m");
        psum::hip_check(hipMalloc(&ia, n * 4), "m");
        psum::hip_check(hipMalloc(&ib, n * 4), "m");
        for (int mode = 0; mode < 3; ++mode) {
            if (mode == 0) trig.set_fast();
            else trig.set_restrict();
            if (mode == 2) trig.set_acceleration_UDF(true);
            psum::handler h{st};
            auto acc = trig.get_access(h);
            hipLaunchKernelGGL(k_detect<decltype(acc)>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, acc, d1,
                               d2, n, ka, na, ia);
            psum::hip_check(hipGetLastError(), "k_detect");
            const void* q1[3] = {d1, d1 + n, d1 + 2 * n};
            const void* q2[3] = {d2, d2 + n, d2 + 2 * n};
            if (psum_hip_mesh_detect(trig.handle(), PSUM_F64, q1, q2, (int64_t)n, kb, nb, ib, st) != PSUM_OK) {
                std::printf("psum_hip_mesh_detect: %s\n", psum_hip_last_error());
                return 2;
            }
            psum::hip_check(hipStreamSynchronize(st), "sync");
            auto hka = host_copy(ka, n), hkb = host_copy(kb, n);
            auto hna = host_copy(na, 3 * n), hnb = host_copy(nb, 3 * n);
            auto hia = host_copy(ia, n), hib = host_copy(ib, n);
            size_t hits = 0;
            for (size_t i = 0; i < n; ++i) hits += hia[i] >= 0;
            const bool same = hia == hib && std::memcmp(hka.data(), hkb.data(), n * 8) == 0 &&
                              std::memcmp(hna.data(), hnb.data(), 3 * n * 8) == 0;
            const char* what[3] = {"detect (fast): accessor == C ABI", "detect (restrict): accessor == C ABI",
                                   "detect (restrict, UDF): accessor == C ABI"};
            report(what[mode], same && hits > 2000, (double)hits);
        }
        for (void* p : {(void*)d1, (void*)d2, (void*)ka, (void*)kb, (void*)na, (void*)nb, (void*)ia, (void*)ib})
            (void)hipFree(p);
    }

    // ---- router: drift + deal in for_each vs psum_hip_mesh_boundary
    {
        Router router(st);
        router.set(meshes, 20, 20, 30);
        router.get_material_set().set_default_temperature(300.0);
        router.get_material_set().set_default_mass(6.67e-27);
        std::mt19937_64 rng(22);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        std::normal_distribution<double> G(0.0, 1.0);
        const size_t n = 6000;
        std::vector<Particle> host(n);
        std::vector<double> pos(3 * n), vel(3 * n);
        std::vector<uint32_t> seed(n);
        for (size_t i = 0; i < n; ++i) {
            Particle& p = host[i];
            auto& x = get<position>(p);
            auto& v = get<velocity>(p);
            x[0] = -0.6 + 1.2 * U(rng);
            x[1] = -0.6 + 1.2 * U(rng);
            x[2] = -3.9 + 6.9 * U(rng);
            for (int c = 0; c < 3; ++c) {
                v[c] = G(rng);
                pos[c * n + i] = x[c];
                vel[c * n + i] = v[c];
            }
            get<random_seed>(p) = seed[i] = (uint32_t)rng();
            get<weight>(p) = (double)i;
        }
        Group group(st, n);
        group.insert(host);
        double* dold = dev_copy(pos);
        double* dpos = dev_copy(pos);
        double* dvel = dev_copy(vel);
        uint32_t* dseed = dev_copy(seed);
        unsigned long long* dhits;
        psum::hip_check(hipMalloc(&dhits, 8), "m");
        // the API side's state, written out by index after each step
        double* apos = dev_copy(pos);
        double* avel = dev_copy(vel);
        uint32_t* aseed = dev_copy(seed);
        const double dt = 0.1;
        bool same = true;
        unsigned long long total = 0;
        for (int step = 0; step < 40 && same; ++step) {
            // C ABI: old = pos; pos += v dt; boundary
            auto hp = host_copy(dpos, 3 * n), hv = host_copy(dvel, 3 * n);
            std::vector<double> np(3 * n);
            for (size_t j = 0; j < 3 * n; ++j) np[j] = hp[j] + hv[j] * dt;
            psum::hip_check(hipMemcpy(dold, hp.data(), 3 * n * 8, hipMemcpyHostToDevice), "copy");
            psum::hip_check(hipMemcpy(dpos, np.data(), 3 * n * 8, hipMemcpyHostToDevice), "copy");
            psum::hip_check(hipMemset(dhits, 0, 8), "memset");
            const void* o3[3] = {dold, dold + n, dold + 2 * n};
            void* p3[3] = {dpos, dpos + n, dpos + 2 * n};
            void* v3[3] = {dvel, dvel + n, dvel + 2 * n};
            if (psum_hip_mesh_boundary(router.get_trigger().handle(), PSUM_F64, o3, p3, v3, dseed, (int64_t)n, 300.0,
                                       6.67e-27, nullptr, dhits, st) != PSUM_OK) {
                std::printf("psum_hip_mesh_boundary: %s\n", psum_hip_last_error());
                return 2;
            }
            // API: the reference step loop shape
            group.for_each([&](psum::handler& h) {
                auto acc = router.get_access(h);
                return [=](Particle& p) {
                    auto& x = get<position>(p);
                    const auto& v = get<velocity>(p);
                    const double o0 = x[0], o1 = x[1], o2 = x[2];
                    for (int c = 0; c < 3; ++c) x[c] = x[c] + v[c] * dt;
                    acc.deal(o0, o1, o2, x[0], x[1], x[2], p);
                };
            });
            // records that are no longer valid are not visited: their x
            // stays all-ones bits, a NaN
            psum::hip_check(hipMemsetAsync(apos, 0xFF, n * 8, st), "memset");
            group.for_each([&](psum::handler&) {
                return [=](Particle& p) {
                    const size_t i = (size_t)get<weight>(p);
                    for (int c = 0; c < 3; ++c) {
                        apos[c * n + i] = get<position>(p)[c];
                        avel[c * n + i] = get<velocity>(p)[c];
                    }
                    aseed[i] = get<random_seed>(p);
                };
            });
            psum::hip_check(hipStreamSynchronize(st), "sync");
            auto cp = host_copy(dpos, 3 * n), cv = host_copy(dvel, 3 * n), ap = host_copy(apos, 3 * n),
                 av = host_copy(avel, 3 * n);
            auto cs = host_copy(dseed, n), as = host_copy(aseed, n);
            total += host_copy(dhits, 1)[0];
            for (size_t i = 0; i < n && same; ++i) {
                if (std::isnan(cp[i])) {
                    // absorbed: the API record is invalid (x NaN) and no longer visited
                    same = std::isnan(ap[i]);
                    if (!same) std::printf("  step %d particle %zu absorbed by the C ABI only\n", step, i);
                    continue;
                }
                for (int c = 0; c < 3; ++c)
                    same = same && std::memcmp(&cp[c * n + i], &ap[c * n + i], 8) == 0 &&
                           std::memcmp(&cv[c * n + i], &av[c * n + i], 8) == 0;
                same = same && cs[i] == as[i];
                if (!same) std::printf("  step %d particle %zu differs\n", step, i);
            }
        }
        report("router: for_each deal == C ABI (40 steps)", same && total > 500, (double)total);
        for (void* p : {(void*)dold, (void*)dpos, (void*)dvel, (void*)dseed, (void*)dhits, (void*)apos, (void*)avel,
                        (void*)aseed})
            (void)hipFree(p);
    }
    psum::hip_check(hipStreamDestroy(st), "stream");
    std::printf(fails ? "FAILED %d\n" : "ALL OK\n", fails);
    return fails ? 1 : 0;
}
