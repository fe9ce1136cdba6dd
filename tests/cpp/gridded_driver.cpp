// Reduced range-dependent scenario (gridded sound speed and bathymetry)
// driven through the C ABI from a flat description written by
// tests/test_msan.py, for the MemorySanitizer build of the CPU emulation.
// Prints the record counts.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "../../include/wq3d.h"

static std::vector<std::vector<double>> arr;
static wq3d_seq seq(int v, int inc, int search) {
    wq3d_seq s{};
    s.size = (int64_t)arr[v].size();
    s.value = arr[v].data();
    s.increment = arr[inc].data();
    s.search = search;
    return s;
}
static void grid(FILE* f, wq3d_grid& g) {
    int nd, it[3], ed[3], data, ax[3][3];
    if (fscanf(f, "%d %d %d %d %d %d %d %d", &nd, &it[0], &it[1], &it[2], &ed[0], &ed[1], &ed[2], &data) != 8) exit(2);
    for (int k = 0; k < 3; ++k)
        if (fscanf(f, "%d %d %d", &ax[k][0], &ax[k][1], &ax[k][2]) != 3) exit(2);
    g.ndims = nd;
    for (int k = 0; k < 3; ++k) {
        g.interp[k] = it[k];
        g.edge_limit[k] = ed[k];
        if (k < nd) g.axis[k] = seq(ax[k][0], ax[k][1], ax[k][2]);
    }
    g.data = arr[data].data();
}
int main(int argc, char** argv) {
    FILE* f = fopen(argv[1], "r");
    FILE* b = fopen(argv[2], "rb");
    int n;
    if (fscanf(f, "%d", &n) != 1) return 2;
    arr.resize(n);
    for (int k = 0; k < n; ++k) {
        long sz;
        if (fscanf(f, "%ld", &sz) != 1) return 2;
        arr[k].resize(sz);
    }
    for (int k = 0; k < n; ++k)
        if (fread(arr[k].data(), 8, arr[k].size(), b) != arr[k].size()) return 3;
    double R;
    if (fscanf(f, "%lf", &R) != 1) return 2;
    wq3d_env_desc env{};
    int atten;
    grid(f, env.profile.grid);
    if (fscanf(f, "%d", &atten) != 1) return 2;
    env.profile.kind = WQ3D_PROFILE_GRID;
    env.profile.atten_kind = atten;
    grid(f, env.bottom.grid);
    env.bottom.kind = WQ3D_BOUNDARY_GRID;
    env.bottom.normal_rho = 1.0;
    if (fscanf(f, "%d %lf %lf %lf %lf %lf", &env.bottom.loss_kind, &env.bottom.loss[0], &env.bottom.loss[1],
               &env.bottom.loss[2], &env.bottom.loss[3], &env.bottom.loss[4]) != 6) return 2;
    env.surface.kind = WQ3D_BOUNDARY_FLAT;
    if (fscanf(f, "%lf %lf %d %lf %lf %lf %lf %lf", &env.surface.rho, &env.surface.normal_rho, &env.surface.loss_kind,
               &env.surface.loss[0], &env.surface.loss[1], &env.surface.loss[2], &env.surface.loss[3],
               &env.surface.loss[4]) != 8) return 2;
    wq3d_fan_desc fd{};
    int v, inc, se;
    if (fscanf(f, "%d %d %d", &v, &inc, &se) != 3) return 2;
    fd.freq = seq(v, inc, se);
    if (fscanf(f, "%d %d %d", &v, &inc, &se) != 3) return 2;
    fd.de = seq(v, inc, se);
    if (fscanf(f, "%d %d %d", &v, &inc, &se) != 3) return 2;
    fd.az = seq(v, inc, se);
    int steps;
    if (fscanf(f, "%lf %lf %lf %lf %d", &fd.src_rho, &fd.src_theta, &fd.src_phi, &fd.time_step, &steps) != 5) return 2;
    int n1, n2, tr, tt, tp;
    if (fscanf(f, "%d %d %d %d %d", &n1, &n2, &tr, &tt, &tp) != 5) return 2;
    fd.num_t1 = n1;
    fd.num_t2 = n2;
    fd.target_rho = arr[tr].data();
    fd.target_theta = arr[tt].data();
    fd.target_phi = arr[tp].data();
    fd.earth_radius = R;
    fd.spreading = WQ3D_SPREADING_HYBRID_GAUSSIAN;
    fd.thresholds = {300.0, 1e-30, 999, 999, 999, 999, 999, 0};
    fd.want_eigenverbs = argc > 3;
    fd.want_reflections = argc > 3;
    wq3d_env* e = nullptr;
    wq3d_fan* fan = nullptr;
    if (wq3d_env_create(&env, &e) || wq3d_fan_create(e, &fd, &fan)) {
        printf("create failed: %s\n", wq3d_last_error());
        return 1;
    }
    for (int s = 0; s < steps; ++s)
        if (wq3d_step(fan, 1)) {
            printf("step %d failed: %s\n", s, wq3d_last_error());
            return 1;
        }
    int64_t nr = 0, nv = 0, nf = 0;
    wq3d_drain_eigenrays(fan, nullptr, nullptr, nullptr, 0, &nr);
    wq3d_drain_eigenverbs(fan, nullptr, nullptr, 0, &nv);
    wq3d_drain_reflections(fan, nullptr, 0, &nf);
    printf("eigenrays %lld eigenverbs %lld reflections %lld\n", (long long)nr, (long long)nv, (long long)nf);
    wq3d_fan_destroy(fan);
    wq3d_env_destroy(e);
    return 0;
}
