// Host-side checks of barnes-hut_amd/host/bh_gen.cpp (generators, CSV reader
// and writer) built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitize_host.py.  Exits 0 and prints "ok" when every check
// holds; a sanitizer report aborts the program.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bh.h"

static int check(bool ok, const char* what)
{
    if (!ok) std::fprintf(stderr, "check failed: %s\n", what);
    return ok ? 0 : 1;
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    int bad = 0;
    for (int kind : {BH_GEN_BELT, BH_GEN_KUIPER, BH_GEN_TROJAN}) {
        const int64_t n = 3000;
        std::vector<bh_vec3> pos(n), vel(n), p2(n), v2(n);
        std::vector<double> mass(n), m2(n);
        std::vector<uint64_t> ids(n);
        bad += check(bh_generate(kind, 7, n, pos.data(), vel.data(), mass.data()) == BH_OK, "generate");
        bool finite = true;
        for (int64_t i = 0; i < n; ++i)
            finite = finite && std::isfinite(pos[i].x) && std::isfinite(pos[i].y) && std::isfinite(pos[i].z) &&
                     std::isfinite(vel[i].x) && std::isfinite(mass[i]) && mass[i] > 0.0;
        bad += check(finite, "finite state");
        const std::string path = dir + "/k" + std::to_string(kind) + ".csv";
        bad += check(bh_write_csv(path.c_str(), n, pos.data(), vel.data(), mass.data()) == BH_OK, "write csv");
        int64_t got = 0;
        bad += check(bh_read_csv(path.c_str(), 0, n, &got, p2.data(), v2.data(), m2.data(), ids.data()) == BH_OK,
                     "read csv");
        bad += check(got == n, "csv count");
        bool same = true;
        for (int64_t i = 0; i < n && i < got; ++i)
            same = same && p2[i].x == pos[i].x && p2[i].y == pos[i].y && p2[i].z == pos[i].z &&
                   v2[i].x == vel[i].x && v2[i].z == vel[i].z && m2[i] == mass[i] && ids[i] == uint64_t(i);
        bad += check(same, "csv round trip bit for bit");
        int64_t limited = 0;
        bad += check(bh_read_csv(path.c_str(), 100, n, &limited, p2.data(), v2.data(), m2.data(), ids.data()) ==
                         BH_OK && limited == 100,
                     "csv body limit");
        int64_t counted = 0;
        bad += check(bh_read_csv(path.c_str(), 0, 0, &counted, nullptr, nullptr, nullptr, nullptr) == BH_OK &&
                         counted == n,
                     "csv count only");
        int64_t small = 0;
        bad += check(bh_read_csv(path.c_str(), 0, 10, &small, p2.data(), v2.data(), m2.data(), ids.data()) ==
                         BH_ERR_OVERFLOW,
                     "csv capacity");
    }
    {   // malformed rows are skipped, a long row is read whole
        const std::string path = dir + "/bad.csv";
        FILE* f = std::fopen(path.c_str(), "w");
        std::fprintf(f, "id,name,class,mass,x,y,z,vx,vy,vz\n");
        std::fprintf(f, "0,a,b,1.0,1,2,3,4,5,6\n");
        std::fprintf(f, "1,a,b,not-a-number,1,2,3,4,5,6\n");
        std::fprintf(f, "2,a,b,2.0,1,2\n");
        std::fprintf(f, "3,%s,b,3.0,7,8,9,1,1,1\n", std::string(9000, 'x').c_str());
        std::fprintf(f, "\n");
        std::fclose(f);
        std::vector<bh_vec3> p(8), v(8);
        std::vector<double> m(8);
        std::vector<uint64_t> id(8);
        int64_t got = 0;
        bad += check(bh_read_csv(path.c_str(), 0, 8, &got, p.data(), v.data(), m.data(), id.data()) == BH_OK,
                     "read malformed csv");
        bad += check(got == 2 && m[0] == 1.0 && m[1] == 3.0 && p[1].x == 7.0, "malformed rows skipped");
        int64_t none = 0;
        bad += check(bh_read_csv((dir + "/missing.csv").c_str(), 0, 8, &none, p.data(), v.data(), m.data(),
                                 id.data()) == BH_ERR_IO,
                     "missing file");
    }
    {
        std::vector<int64_t> sizes(64);
        bad += check(bh_generate_ensemble_sizes(0, 64, 10000, 200000, sizes.data()) == BH_OK, "ensemble sizes");
        bool in = true;
        for (int64_t s : sizes) in = in && s >= 10000 && s <= 200000;
        bad += check(in, "ensemble sizes in range");
        bad += check(bh_generate(BH_GEN_BELT, 0, -1, nullptr, nullptr, nullptr) != BH_OK, "negative count");
    }
    if (bad) return 1;
    std::printf("ok\n");
    return 0;
}
