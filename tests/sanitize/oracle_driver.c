/* Runs the oracle's step loop (oracle/bh_oracle.c, or_simulate) on a state
 * read from a binary file and writes the final state, built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_sanitize_host.py, which compares the output with the oracle
 * library's own run bit for bit.
 * Input: int64 n, int64 steps, double dt, double theta, then n x (pos 3,
 * vel 3, mass) doubles.  Output: n x (pos 3, vel 3, acc 3) doubles and n ids. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int or_simulate(int64_t n, double *pos, double *vel, double *mass, uint64_t *ids, double *acc, int64_t nsteps,
                double t0, double dt, double theta, double G, double soft2);

int main(int argc, char **argv)
{
    if (argc < 4) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t n = 0, steps = 0;
    double dt = 0, theta = 0, G = 0;
    if (fread(&n, 8, 1, f) != 1 || fread(&steps, 8, 1, f) != 1 || fread(&dt, 8, 1, f) != 1 ||
        fread(&theta, 8, 1, f) != 1 || fread(&G, 8, 1, f) != 1 || n <= 0)
        return 2;
    double *pos = malloc(sizeof(double) * 3 * n), *vel = malloc(sizeof(double) * 3 * n);
    double *mass = malloc(sizeof(double) * n), *acc = calloc(3 * (size_t)n, sizeof(double));
    uint64_t *ids = malloc(sizeof(uint64_t) * n);
    if (!pos || !vel || !mass || !acc || !ids) return 3;
    if (fread(pos, 8, 3 * n, f) != (size_t)(3 * n) || fread(vel, 8, 3 * n, f) != (size_t)(3 * n) ||
        fread(mass, 8, n, f) != (size_t)n)
        return 2;
    fclose(f);
    for (int64_t i = 0; i < n; ++i) ids[i] = (uint64_t)i;
    const int rc = or_simulate(n, pos, vel, mass, ids, acc, steps, 0.0, dt, theta, G, 0.0);
    if (rc) return 4;
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    fwrite(pos, 8, 3 * n, o);
    fwrite(vel, 8, 3 * n, o);
    fwrite(acc, 8, 3 * n, o);
    fwrite(ids, 8, n, o);
    fclose(o);
    printf("%s\n", argv[3]);
    free(pos);
    free(vel);
    free(mass);
    free(acc);
    free(ids);
    return 0;
}
