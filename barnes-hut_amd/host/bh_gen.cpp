// bh_gen.cpp -- deterministic synthetic heliocentric systems and the
// reference's state-vector CSV format.
//
// Generator.  A central mass at the origin, eight planets on circular
// coplanar orbits, and asteroid populations drawn from orbital elements and
// converted Kepler -> Cartesian with the textbook two-body relations that the
// reference uses for its scenario files (src/kepler_to_cartesian.cpp:33-91:
// Newton iteration on Kepler's equation from E0 = M with |dE| < 1e-8 and at
// most 100 iterations, true anomaly from the half-angle atan2 form, then the
// perifocal -> inertial rotation by (Omega, i, omega)).  Units: AU, kg, day.
// The random stream is xoshiro256** seeded through splitmix64, with uniform
// doubles from the top 53 bits, so a (kind, seed, n) triple always gives the
// same bodies on the same libm.
//
// CSV.  Header line, then `id,name,class,mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z`
// rows (src/io.cpp:16-81): the first three columns are ignored on read, ids
// are assigned in row order, malformed rows are skipped, and a positive body
// limit stops reading early.  Values are written with 17 significant digits so
// a write/read round trip is exact.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/bh.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg = kPi / 180.0;

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t z = seed;
        for (int k = 0; k < 4; ++k) {
            // splitmix64 step
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            s[k] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        // xoshiro256** (Blackman & Vigna)
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }   // [0,1)
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    // Rayleigh(sigma) by inversion of F(x) = 1 - exp(-x^2 / (2 sigma^2))
    double rayleigh(double sigma) { return sigma * std::sqrt(-2.0 * std::log1p(-uniform())); }
    // standard normal, Box-Muller (cosine branch only, one draw per call)
    double normal()
    {
        double u1 = 1.0 - uniform();   // (0,1]
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
    double log_uniform(double lo, double hi)
    {
        return std::exp(std::log(lo) + (std::log(hi) - std::log(lo)) * uniform());
    }
};

struct Elements {
    double a, e, i, node, peri, M;
};

// Two-body state from elements about a central parameter mu (AU^3/day^2).
void kepler_to_state(const Elements& el, double mu, bh_vec3* x, bh_vec3* v)
{
    double E = el.M;
    for (int it = 0; it < 100; ++it) {
        double f = E - el.e * std::sin(E) - el.M;
        double fp = 1.0 - el.e * std::cos(E);
        double dE = -f / fp;
        E += dE;
        if (std::fabs(dE) < 1e-8) break;
    }
    double nu = 2.0 * std::atan2(std::sqrt(1.0 + el.e) * std::sin(0.5 * E),
                                 std::sqrt(1.0 - el.e) * std::cos(0.5 * E));
    double r = el.a * (1.0 - el.e * std::cos(E));
    double px = r * std::cos(nu), py = r * std::sin(nu);
    double h = std::sqrt(mu * el.a) / r;
    double qx = -h * std::sin(E), qy = h * std::sqrt(1.0 - el.e * el.e) * std::cos(E);
    double cO = std::cos(el.node), sO = std::sin(el.node);
    double ci = std::cos(el.i), si = std::sin(el.i);
    double cw = std::cos(el.peri), sw = std::sin(el.peri);
    // columns of the perifocal -> inertial rotation R3(-Omega) R1(-i) R3(-omega)
    double P[3] = {cO * cw - sO * sw * ci, sO * cw + cO * sw * ci, sw * si};
    double Q[3] = {-cO * sw - sO * cw * ci, -sO * sw + cO * cw * ci, cw * si};
    x->x = P[0] * px + Q[0] * py;
    x->y = P[1] * px + Q[1] * py;
    x->z = P[2] * px + Q[2] * py;
    v->x = P[0] * qx + Q[0] * qy;
    v->y = P[1] * qx + Q[1] * qy;
    v->z = P[2] * qx + Q[2] * qy;
}

const double kPlanetA[8] = {0.39, 0.72, 1.0, 1.52, 5.2, 9.54, 19.2, 30.1};
const double kPlanetM[8] = {3.30e23, 4.87e24, 5.97e24, 6.42e23, 1.90e27, 5.68e26, 8.68e25, 1.02e26};

void draw_belt_like(Rng& rng, double mu, double amin, double amax, double emax, double isig,
                    bh_vec3* x, bh_vec3* v, double* m)
{
    Elements el;
    el.a = rng.uniform(amin, amax);
    el.e = rng.uniform(0.0, emax);
    el.i = rng.rayleigh(isig);
    el.node = rng.uniform(0.0, 2.0 * kPi);
    el.peri = rng.uniform(0.0, 2.0 * kPi);
    el.M = rng.uniform(0.0, 2.0 * kPi);
    *m = rng.log_uniform(1e12, 1e20);
    kepler_to_state(el, mu, x, v);
}

}  // namespace

extern "C" {

int bh_generate(int kind, uint64_t seed, int64_t n, bh_vec3* pos, bh_vec3* vel, double* mass)
{
    if (n < 9 || !pos || !vel || !mass) return BH_ERR_ARG;
    if (kind != BH_GEN_BELT && kind != BH_GEN_KUIPER && kind != BH_GEN_TROJAN) return BH_ERR_ARG;
    Rng rng(seed ^ (0x5eedULL << 48) ^ static_cast<uint64_t>(kind));
    const double mu = BH_G * BH_GEN_CENTRAL_MASS;
    pos[0] = {0, 0, 0};
    vel[0] = {0, 0, 0};
    mass[0] = BH_GEN_CENTRAL_MASS;
    double jupiter_phase = 0.0;
    for (int p = 0; p < 8; ++p) {
        double phase = rng.uniform(0.0, 2.0 * kPi);
        double a = kPlanetA[p];
        double vc = std::sqrt(mu / a);
        pos[1 + p] = {a * std::cos(phase), a * std::sin(phase), 0.0};
        vel[1 + p] = {-vc * std::sin(phase), vc * std::cos(phase), 0.0};
        mass[1 + p] = kPlanetM[p];
        if (p == 4) jupiter_phase = phase;
    }
    const int64_t na = n - 9;
    int64_t n_kuiper = 0, n_trojan = 0, n_hilda = 0;
    if (kind == BH_GEN_KUIPER) n_kuiper = (na * 3) / 10;
    if (kind == BH_GEN_TROJAN) {
        n_trojan = (na * 15) / 100;
        n_hilda = na / 10;
    }
    const int64_t n_belt = na - n_kuiper - 2 * n_trojan - n_hilda;
    int64_t k = 9;
    for (int64_t j = 0; j < n_belt; ++j, ++k)
        draw_belt_like(rng, mu, 2.1, 3.3, 0.3, 5.0 * kDeg, &pos[k], &vel[k], &mass[k]);
    for (int64_t j = 0; j < n_kuiper; ++j, ++k)
        draw_belt_like(rng, mu, 30.0, 50.0, 0.2, 10.0 * kDeg, &pos[k], &vel[k], &mass[k]);
    for (int side = 0; side < 2 && n_trojan > 0; ++side) {
        const double lam = jupiter_phase + (side == 0 ? 60.0 : -60.0) * kDeg;
        const double cx = kPlanetA[4] * std::cos(lam), cy = kPlanetA[4] * std::sin(lam);
        for (int64_t j = 0; j < n_trojan; ++j, ++k) {
            double x = cx + 0.15 * rng.normal();
            double y = cy + 0.15 * rng.normal();
            double z = 0.15 * rng.normal();
            double rxy = std::sqrt(x * x + y * y);
            double vc = std::sqrt(mu / rxy);
            pos[k] = {x, y, z};
            vel[k] = {-vc * y / rxy, vc * x / rxy, 0.0};
            mass[k] = rng.log_uniform(1e12, 1e20);
        }
    }
    for (int64_t j = 0; j < n_hilda; ++j, ++k)
        draw_belt_like(rng, mu, 3.7, 4.2, 0.3, 5.0 * kDeg, &pos[k], &vel[k], &mass[k]);
    return k == n ? BH_OK : BH_ERR_STATE;
}

int bh_generate_ensemble_sizes(uint64_t seed, int64_t count, int64_t nmin, int64_t nmax, int64_t* sizes)
{
    if (count < 0 || nmin < 9 || nmax < nmin || (count > 0 && !sizes)) return BH_ERR_ARG;
    Rng rng(seed ^ 0xe45e3b1eULL);
    for (int64_t s = 0; s < count; ++s) {
        double v = rng.log_uniform(static_cast<double>(nmin), static_cast<double>(nmax));
        int64_t q = static_cast<int64_t>(std::llround(v));
        sizes[s] = q < nmin ? nmin : (q > nmax ? nmax : q);
    }
    return BH_OK;
}

int bh_write_csv(const char* path, int64_t n, const bh_vec3* pos, const bh_vec3* vel, const double* mass)
{
    if (!path || n < 0 || (n > 0 && (!pos || !vel || !mass))) return BH_ERR_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return BH_ERR_IO;
    std::fprintf(f, "id,name,class,mass,pos_x,pos_y,pos_z,vel_x,vel_y,vel_z\n");
    for (int64_t i = 0; i < n; ++i) {
        const char* cls = i == 0 ? "STA" : (i < 9 ? "PLA" : "AST");
        std::fprintf(f, "%lld,b%lld,%s,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n",
                     static_cast<long long>(i), static_cast<long long>(i), cls, mass[i], pos[i].x,
                     pos[i].y, pos[i].z, vel[i].x, vel[i].y, vel[i].z);
    }
    return std::fclose(f) == 0 ? BH_OK : BH_ERR_IO;
}

namespace {

// strtod over one comma-separated field; false when the field is not a full
// number (std::stod in the reference throws for a field without a leading
// number; trailing characters after a number are accepted like std::stod).
bool parse_field(const char*& p, double* out)
{
    const char* end = std::strchr(p, ',');
    std::string field = end ? std::string(p, end) : std::string(p);
    p = end ? end + 1 : p + std::strlen(p);
    const char* s = field.c_str();
    char* e = nullptr;
    errno = 0;
    double v = std::strtod(s, &e);
    if (e == s || errno == ERANGE) return false;
    *out = v;
    return true;
}

bool skip_field(const char*& p)
{
    const char* end = std::strchr(p, ',');
    if (!end) {
        p += std::strlen(p);
        return true;   // std::getline on an exhausted stream leaves the field empty
    }
    p = end + 1;
    return true;
}

}  // namespace

int bh_read_csv(const char* path, int64_t body_limit, int64_t capacity, int64_t* n_out, bh_vec3* pos,
                bh_vec3* vel, double* mass, uint64_t* ids)
{
    if (!path || !n_out || capacity < 0) return BH_ERR_ARG;
    FILE* f = std::fopen(path, "r");
    if (!f) return BH_ERR_IO;
    std::string line;
    char buf[4096];
    bool header = true;
    int64_t n = 0;
    int rc = BH_OK;
    for (;;) {
        line.clear();
        bool got = false;
        while (std::fgets(buf, sizeof buf, f)) {
            got = true;
            line += buf;
            if (!line.empty() && line.back() == '\n') break;
        }
        if (!got) break;
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
        if (header) {
            header = false;
            continue;
        }
        if (body_limit > 0 && n >= body_limit) break;
        const char* p = line.c_str();
        double v[7];
        skip_field(p);
        skip_field(p);
        skip_field(p);
        bool ok = true;
        for (int c = 0; c < 7 && ok; ++c) ok = parse_field(p, &v[c]);
        if (!ok) {
            std::fprintf(stderr, "Warning: Could not parse line, skipping: \"%s\"\n", line.c_str());
            continue;
        }
        if (pos && vel && mass && n < capacity) {
            mass[n] = v[0];
            pos[n] = {v[1], v[2], v[3]};
            vel[n] = {v[4], v[5], v[6]};
            if (ids) ids[n] = static_cast<uint64_t>(n);
        } else if (pos || vel || mass) {
            rc = BH_ERR_OVERFLOW;
        }
        ++n;
    }
    std::fclose(f);
    *n_out = n;
    return rc;
}

}  // extern "C"
