// bh_dropin.cpp -- the reference's step loop (src/main.cpp:191-281) on the
// host, with the engine as a per-call drop-in for its `--fc tree` force path
// (src/main.cpp:212-263: box, keys, sortBodiesByMortonKey, buildOctreeBottomUp,
// exchangeFullTrees, computeAccelerations).
//
// The host owns the bodies in std::vector<Vec3> as the reference does; the
// half kicks and the drift stay on the host.  Each step uploads the drifted
// state, lets the engine sort it into Morton order and compute the
// accelerations, and reads the permuted bodies back, so the host arrays
// follow the same persistent permutation as the reference's
// sortBodiesByMortonKey (include/utility.h:23-52) and the accelerations line
// up with the velocities they kick.
//
//   bh_dropin CSV DT TEND THETA OUT.csv
// reads a state-vector CSV (src/io.cpp:16-81), runs `for (t = 0; t < TEND;
// t += DT)` (times as "1h", "2d", ...), and writes id, position, velocity and
// acceleration of every body in the final order with 17 significant digits.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/bh.h"

// include/body.h:4-6
struct Vec3 {
    double x, y, z;
};
static_assert(sizeof(Vec3) == sizeof(bh_vec3), "Vec3 and bh_vec3 share one layout");

namespace {

int fail(const char* what, bh_ctx* ctx)
{
    std::fprintf(stderr, "Error: %s: %s\n", what, bh_last_error(ctx));
    return 1;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s CSV DT TEND THETA OUT.csv\n", argv[0]);
        return 2;
    }
    double dt = 0.0, t_end = 0.0;
    if (bh_parse_time(argv[2], &dt) != BH_OK || bh_parse_time(argv[3], &t_end) != BH_OK || !(dt > 0.0)) {
        std::fprintf(stderr, "Error: times are a number and a unit h, d or y\n");
        return 2;
    }
    const double theta = std::atof(argv[4]);
    const double G = BH_G;
    int64_t N = 0;
    if (bh_read_csv(argv[1], -1, 0, &N, nullptr, nullptr, nullptr, nullptr) != BH_OK) return fail("read", nullptr);
    std::vector<Vec3> local_pos(N), local_vel(N), local_acc(N, Vec3{0.0, 0.0, 0.0});
    std::vector<double> local_mass(N);
    std::vector<uint64_t> local_ids(N);
    int64_t got = 0;
    if (bh_read_csv(argv[1], -1, N, &got, reinterpret_cast<bh_vec3*>(local_pos.data()),
                    reinterpret_cast<bh_vec3*>(local_vel.data()), local_mass.data(), local_ids.data()) != BH_OK)
        return fail("read", nullptr);

    // [sketch begin]
    bh_config cfg;
    bh_config_default(&cfg);          // full tree, PLAIN, theta 1.05, G, soft2 = 0
    cfg.theta = theta;
    bh_ctx* ctx = nullptr;
    if (bh_create(&ctx, &cfg) != BH_OK) return fail("bh_create", nullptr);

    double t = 0.0;
    for (int step = 0; t < t_end; ++step, t += dt) {
        // 1. half kick (main.cpp:193-200)
        if (t > 0) {
            for (int64_t i = 0; i < N; ++i) {
                local_vel[i].x += 0.5 * dt * local_acc[i].x;
                local_vel[i].y += 0.5 * dt * local_acc[i].y;
                local_vel[i].z += 0.5 * dt * local_acc[i].z;
            }
        }
        // 2. drift (main.cpp:202-208)
        for (int64_t i = 0; i < N; ++i) {
            local_pos[i].x += dt * local_vel[i].x;
            local_pos[i].y += dt * local_vel[i].y;
            local_pos[i].z += dt * local_vel[i].z;
        }
        // 3-7. box, keys, sort, build, exchange, walk (main.cpp:212-263):
        // upload the drifted state, evaluate, and read the bodies back in
        // the new Morton order, which the next kick and drift keep using
        if (bh_set_bodies(ctx, N, reinterpret_cast<const bh_vec3*>(local_pos.data()),
                          reinterpret_cast<const bh_vec3*>(local_vel.data()), local_mass.data(),
                          local_ids.data()) != BH_OK ||
            bh_compute_accelerations(ctx, BH_FORM_PLAIN, theta, G, 0.0, nullptr) != BH_OK ||
            bh_get_bodies(ctx, reinterpret_cast<bh_vec3*>(local_pos.data()),
                          reinterpret_cast<bh_vec3*>(local_vel.data()), local_mass.data(), local_ids.data(),
                          reinterpret_cast<bh_vec3*>(local_acc.data())) != BH_OK)
            return fail("force evaluation", ctx);
        // 8. second half kick (main.cpp:276-281)
        for (int64_t i = 0; i < N; ++i) {
            local_vel[i].x += 0.5 * dt * local_acc[i].x;
            local_vel[i].y += 0.5 * dt * local_acc[i].y;
            local_vel[i].z += 0.5 * dt * local_acc[i].z;
        }
    }
    bh_destroy(ctx);
    // [sketch end]

    std::FILE* f = std::fopen(argv[5], "w");
    if (!f) {
        std::fprintf(stderr, "Error: cannot write %s\n", argv[5]);
        return 1;
    }
    std::fprintf(f, "id,x,y,z,vx,vy,vz,ax,ay,az\n");
    for (int64_t i = 0; i < N; ++i)
        std::fprintf(f, "%llu,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n",
                     static_cast<unsigned long long>(local_ids[i]), local_pos[i].x, local_pos[i].y, local_pos[i].z,
                     local_vel[i].x, local_vel[i].y, local_vel[i].z, local_acc[i].x, local_acc[i].y,
                     local_acc[i].z);
    std::fclose(f);
    std::printf("bh_dropin: %lld bodies, t = %.17g\n", static_cast<long long>(N), t);
    return 0;
}
