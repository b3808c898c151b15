/*
 * bh_oracle.c -- CPU restatement of the reference's single-rank `--fc tree`
 * force path and step loop (YonatanGM/Barnes-Hut).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (libbh.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  The arithmetic
 * is written in the reference's operation order and compiled with
 * -O2 -ffp-contract=off and no fast-math flags, so it reproduces the strict
 * (RelWithDebInfo-style) reference build bit for bit.  Pins (tests/):
 *   - or_morton_keys:   known answers (unit axes, all-ones key), the spread /
 *                       compact round trip, and oracle/_ref fixtures.
 *   - or_sort:          sortedness + stability + oracle/_ref fixtures.
 *   - or_tree_*:        closed-form topology (unique prefixes per depth),
 *                       brute-force mass / COM on small inputs, oracle/_ref
 *                       fixtures (bitwise COM after the merge round trip).
 *   - or_walk:          theta=0 equals a direct sum over other leaves,
 *                       theta=1e6 equals the root monopole, two/three-body
 *                       closed forms, oracle/_ref fixtures (bitwise).
 *   - or_direct:        closed-form two/three-body accelerations.
 *   - or_simulate:      one-step closed form and oracle/_ref trajectories.
 *
 * Tie order of duplicate Morton keys: the reference uses std::sort (unstable,
 * include/utility.h:35-38); this restatement uses a stable order (ties by
 * current index).  Both agree whenever keys are distinct.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX_DEPTH 21
#define OR_OK 0
#define OR_ERR_ARG -1
#define OR_ERR_NOMEM -2

/* ---------------------------------------------------------------- bbox --- */

/* bounding_box.cpp:23-44 (compute_local_bbox) and :46-52 (one rank: the
 * allreduce is the identity).  bb = {minx,miny,minz,maxx,maxy,maxz}.  An empty
 * input leaves the +inf/-inf box of bounding_box.h:11-22. */
int or_bbox(int64_t n, const double *pos, double *bb)
{
    if (n < 0 || (n > 0 && !pos) || !bb) return OR_ERR_ARG;
    if (n == 0) {
        bb[0] = bb[1] = bb[2] = INFINITY;
        bb[3] = bb[4] = bb[5] = -INFINITY;
        return OR_OK;
    }
    for (int a = 0; a < 3; ++a) { bb[a] = pos[a]; bb[3 + a] = pos[a]; }
    for (int64_t i = 1; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            double v = pos[3 * i + a];
            /* std::min(a,b) returns a unless b < a; std::max returns a unless a < b */
            if (v < bb[a]) bb[a] = v;
            if (bb[3 + a] < v) bb[3 + a] = v;
        }
    }
    return OR_OK;
}

/* ------------------------------------------------------------- morton --- */

/* morton_keys.cpp:11-19: bit b of v goes to bit 3b.  Written as a plain bit
 * loop so that it does not share the reference's mask sequence. */
uint64_t or_spread21(uint32_t v)
{
    uint64_t out = 0;
    for (int b = 0; b < 21; ++b) out |= (uint64_t)((v >> b) & 1u) << (3 * b);
    return out;
}

uint32_t or_compact21(uint64_t m)
{
    uint32_t out = 0;
    for (int b = 0; b < 21; ++b) out |= (uint32_t)((m >> (3 * b)) & 1u) << b;
    return out;
}

/* morton_keys.cpp:22-25: x in bit 0, y in bit 1, z in bit 2. */
uint64_t or_morton63(uint32_t xi, uint32_t yi, uint32_t zi)
{
    return or_spread21(xi) | (or_spread21(yi) << 1) | (or_spread21(zi) << 2);
}

/* morton_keys.cpp:28-32 */
uint64_t or_prefix(uint64_t code, int depth)
{
    if (depth == 0) return 0;
    return code & (~0ULL << (63 - 3 * depth));
}

/* morton_keys.cpp:75-83: normalise with clamping to [0, nextafter(1,0)]. */
static double or_norm(double v, double lo, double span)
{
    double t = (v - lo) / span;
    if (t <= 0.0) return 0.0;
    if (t >= 1.0) return nextafter(1.0, 0.0);
    return t;
}

/* morton_keys.cpp:65-98 (generateMortonCodes). */
int or_morton_keys(int64_t n, const double *pos, const double *bb, uint64_t *keys)
{
    if (n < 0 || (n > 0 && (!pos || !keys)) || !bb) return OR_ERR_ARG;
    double span[3];
    for (int a = 0; a < 3; ++a)
        span[a] = (bb[3 + a] == bb[a]) ? 1.0 : (bb[3 + a] - bb[a]);
    const uint32_t Q = 1u << 21;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            double t = or_norm(pos[3 * i + a], bb[a], span[a]);
            q[a] = (uint32_t)(t * Q);
        }
        keys[i] = or_morton63(q[0], q[1], q[2]);
    }
    return OR_OK;
}

/* --------------------------------------------------------------- sort --- */

static void or_merge_sort_idx(int64_t *idx, int64_t *tmp, int64_t n, const uint64_t *k)
{
    /* bottom-up stable merge sort of indices by key */
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) tmp[o++] = (k[idx[b]] < k[idx[a]]) ? idx[b++] : idx[a++];
            while (a < mid) tmp[o++] = idx[a++];
            while (b < hi) tmp[o++] = idx[b++];
        }
        memcpy(idx, tmp, (size_t)n * sizeof(int64_t));
    }
}

/* include/utility.h:23-52 (sortBodiesByMortonKey): argsort by key, then
 * permute keys / pos / mass / vel / ids.  Any array argument may be NULL.
 * perm_out (optional) receives the gather index: new[i] = old[perm_out[i]]. */
int or_sort(int64_t n, uint64_t *keys, double *pos, double *mass, double *vel,
            uint64_t *ids, int64_t *perm_out)
{
    if (n < 0 || (n > 0 && !keys)) return OR_ERR_ARG;
    if (n == 0) return OR_OK;
    int64_t *idx = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *tmp = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    void *buf = malloc((size_t)n * 3 * sizeof(double));
    if (!idx || !tmp || !buf) { free(idx); free(tmp); free(buf); return OR_ERR_NOMEM; }
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    or_merge_sort_idx(idx, tmp, n, keys);
#define OR_PERMUTE(arr, T, W)                                                   \
    if (arr) {                                                                  \
        T *b_ = (T *)buf;                                                       \
        for (int64_t i = 0; i < n; ++i)                                         \
            for (int w_ = 0; w_ < (W); ++w_) b_[(W) * i + w_] = (arr)[(W) * idx[i] + w_]; \
        memcpy((arr), b_, (size_t)n * (W) * sizeof(T));                         \
    }
    OR_PERMUTE(pos, double, 3)
    OR_PERMUTE(mass, double, 1)
    OR_PERMUTE(vel, double, 3)
    OR_PERMUTE(ids, uint64_t, 1)
    OR_PERMUTE(keys, uint64_t, 1)
#undef OR_PERMUTE
    if (perm_out) memcpy(perm_out, idx, (size_t)n * sizeof(int64_t));
    free(idx); free(tmp); free(buf);
    return OR_OK;
}

/* --------------------------------------------------------------- tree --- */

/* A node of the reference's hash-map octree (include/linear_octree.h:15-57):
 * key (prefix, depth), value (mass, com). */
typedef struct or_tree {
    int64_t n_nodes, cap_nodes;
    uint64_t *prefix;
    uint8_t *depth;
    double *mass, *cx, *cy, *cz;
    /* open-addressing index: slot -> node+1 (0 = empty) */
    int64_t *slot;
    uint64_t mask;
} or_tree;

static uint64_t or_hash(uint64_t prefix, int depth)
{
    /* MurmurHash3 64-bit finaliser over prefix + depth in the free top bit area */
    uint64_t h = prefix ^ ((uint64_t)depth * 0xff51afd7ed558ccdULL);
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ULL;
    h ^= h >> 33;
    return h;
}

static int64_t or_find(const or_tree *t, uint64_t prefix, int depth)
{
    uint64_t s = or_hash(prefix, depth) & t->mask;
    for (;;) {
        int64_t v = t->slot[s];
        if (v == 0) return -1;
        if (t->prefix[v - 1] == prefix && t->depth[v - 1] == (uint8_t)depth) return v - 1;
        s = (s + 1) & t->mask;
    }
}

/* find-or-create with zero-initialised value (OctreeMap::operator[]) */
static int64_t or_get(or_tree *t, uint64_t prefix, int depth)
{
    uint64_t s = or_hash(prefix, depth) & t->mask;
    for (;;) {
        int64_t v = t->slot[s];
        if (v == 0) break;
        if (t->prefix[v - 1] == prefix && t->depth[v - 1] == (uint8_t)depth) return v - 1;
        s = (s + 1) & t->mask;
    }
    int64_t id = t->n_nodes++;
    t->prefix[id] = prefix;
    t->depth[id] = (uint8_t)depth;
    t->mass[id] = 0.0; t->cx[id] = 0.0; t->cy[id] = 0.0; t->cz[id] = 0.0;
    t->slot[s] = id + 1;
    return id;
}

void or_tree_free(or_tree *t)
{
    if (!t) return;
    free(t->prefix); free(t->depth); free(t->mass); free(t->cx); free(t->cy); free(t->cz);
    free(t->slot); free(t);
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Unique in place over a sorted array; returns the new length. */
static int64_t or_unique(uint64_t *v, int64_t n)
{
    if (n == 0) return 0;
    int64_t o = 1;
    for (int64_t i = 1; i < n; ++i)
        if (v[i] != v[o - 1]) v[o++] = v[i];
    return o;
}

/* linear_octree.cpp:13-107 (buildOctreeBottomUp) followed by the one-rank
 * full-tree exchange, exchange.cpp:9-29 -> linear_octree.cpp:169-204
 * (serializeTreeToRecords, then mergeRecordsIntoTree into an empty map).
 * keys/pos/mass are in sorted (Morton) order.  Returns NULL on error. */
or_tree *or_tree_build(int64_t n, const uint64_t *keys, const double *pos, const double *mass)
{
    if (n < 0 || (n > 0 && (!keys || !pos || !mass))) return NULL;
    or_tree *t = (or_tree *)calloc(1, sizeof(or_tree));
    if (!t) return NULL;
    int64_t cap = 22 * (n > 0 ? n : 1) + 1;
    t->cap_nodes = cap;
    uint64_t sc = 1;
    while (sc < (uint64_t)(2 * cap)) sc <<= 1;
    t->mask = sc - 1;
    t->prefix = (uint64_t *)malloc((size_t)cap * sizeof(uint64_t));
    t->depth = (uint8_t *)malloc((size_t)cap);
    t->mass = (double *)malloc((size_t)cap * sizeof(double));
    t->cx = (double *)malloc((size_t)cap * sizeof(double));
    t->cy = (double *)malloc((size_t)cap * sizeof(double));
    t->cz = (double *)malloc((size_t)cap * sizeof(double));
    t->slot = (int64_t *)calloc(sc, sizeof(int64_t));
    uint64_t *cur = (uint64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    uint64_t *par = (uint64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint64_t));
    if (!t->prefix || !t->depth || !t->mass || !t->cx || !t->cy || !t->cz || !t->slot || !cur || !par) {
        free(cur); free(par); or_tree_free(t); return NULL;
    }
    if (n == 0) { free(cur); free(par); return t; }

    /* 1. leaves (linear_octree.cpp:28-38): accumulate m and x*m in body order */
    for (int64_t i = 0; i < n; ++i) {
        int64_t id = or_get(t, keys[i], OR_MAX_DEPTH);
        double m = mass[i];
        t->mass[id] += m;
        t->cx[id] += pos[3 * i + 0] * m;
        t->cy[id] += pos[3 * i + 1] * m;
        t->cz[id] += pos[3 * i + 2] * m;
        cur[i] = keys[i];
    }
    qsort(cur, (size_t)n, sizeof(uint64_t), cmp_u64);
    int64_t ncur = or_unique(cur, n);

    /* 2. parents level by level (linear_octree.cpp:49-91): each new parent
     * starts at zero and adds its existing children in octant order 0..7 */
    for (int d = OR_MAX_DEPTH - 1; d >= 0; --d) {
        if (ncur == 0) break;
        for (int64_t j = 0; j < ncur; ++j) par[j] = or_prefix(cur[j], d);
        int64_t npar = or_unique(par, ncur); /* children sorted => parents sorted */
        for (int64_t j = 0; j < npar; ++j) {
            int64_t p = or_get(t, par[j], d);
            for (int oc = 0; oc < 8; ++oc) {
                uint64_t cp = par[j] | ((uint64_t)oc << (63 - 3 * (d + 1)));
                int64_t c = or_find(t, cp, d + 1);
                if (c < 0) continue;
                t->mass[p] += t->mass[c];
                t->cx[p] += t->cx[c];
                t->cy[p] += t->cy[c];
                t->cz[p] += t->cz[c];
            }
        }
        uint64_t *sw = cur; cur = par; par = sw;
        ncur = npar;
    }

    /* 3. finalise (linear_octree.cpp:95-104) */
    for (int64_t k = 0; k < t->n_nodes; ++k) {
        double m = t->mass[k];
        if (m > 0) { t->cx[k] /= m; t->cy[k] /= m; t->cz[k] /= m; }
    }

    /* 4. one-rank exchange round trip (linear_octree.cpp:178-204 into an
     * empty map): m0 = 0, com0 = 0 */
    for (int64_t k = 0; k < t->n_nodes; ++k) {
        double m0 = 0.0, c0 = 0.0;
        double m1 = t->mass[k];
        double total = m0 + m1;
        if (total == 0) {
            t->mass[k] = 0.0; t->cx[k] = 0.0; t->cy[k] = 0.0; t->cz[k] = 0.0;
            continue;
        }
        t->cx[k] = (m0 * c0 + m1 * t->cx[k]) / total;
        t->cy[k] = (m0 * c0 + m1 * t->cy[k]) / total;
        t->cz[k] = (m0 * c0 + m1 * t->cz[k]) / total;
        t->mass[k] = total;
    }
    free(cur); free(par);
    return t;
}

int64_t or_tree_count(const or_tree *t) { return t ? t->n_nodes : -1; }

typedef struct { uint64_t prefix; uint8_t depth; int64_t id; } or_nk;

static int cmp_nk(const void *a, const void *b)
{
    const or_nk *x = (const or_nk *)a, *y = (const or_nk *)b;
    if (x->depth != y->depth) return (x->depth > y->depth) - (x->depth < y->depth);
    return (x->prefix > y->prefix) - (x->prefix < y->prefix);
}

/* Export every node sorted by (depth, prefix): the topology set plus the
 * final (mass, com) the walk reads.  Arrays sized or_tree_count(). */
int or_tree_export(const or_tree *t, uint64_t *prefix, uint8_t *depth, double *mass, double *com)
{
    if (!t) return OR_ERR_ARG;
    int64_t n = t->n_nodes;
    or_nk *v = (or_nk *)malloc((size_t)(n > 0 ? n : 1) * sizeof(or_nk));
    if (!v) return OR_ERR_NOMEM;
    for (int64_t k = 0; k < n; ++k) { v[k].prefix = t->prefix[k]; v[k].depth = t->depth[k]; v[k].id = k; }
    qsort(v, (size_t)n, sizeof(or_nk), cmp_nk);
    for (int64_t k = 0; k < n; ++k) {
        int64_t id = v[k].id;
        if (prefix) prefix[k] = t->prefix[id];
        if (depth) depth[k] = t->depth[id];
        if (mass) mass[k] = t->mass[id];
        if (com) { com[3 * k] = t->cx[id]; com[3 * k + 1] = t->cy[id]; com[3 * k + 2] = t->cz[id]; }
    }
    free(v);
    return OR_OK;
}

/* --------------------------------------------------------------- walk --- */

/* traversal.cpp:13-103 (computeAccelerations) with the reference's stack
 * DFS: children pushed in octant order 0..7, popped 7..0.  keys/pos in
 * sorted order; acc out per body.  n_accept / n_open (optional) count per
 * body the accepted nodes and the opened nodes of the full-chain tree. */
int or_walk(const or_tree *t, int64_t n, const uint64_t *keys, const double *pos,
            const double *bb, double theta, double G, double soft2, double *acc,
            int64_t *n_accept, int64_t *n_open)
{
    if (!t || n < 0 || (n > 0 && (!keys || !pos || !acc)) || !bb) return OR_ERR_ARG;
    const double theta2 = theta * theta;
    double dx = bb[3] - bb[0], dy = bb[4] - bb[1], dz = bb[5] - bb[2];
    /* std::max({dx,dy,dz}) keeps the first of equal maxima */
    double L = dx;
    if (L < dy) L = dy;
    if (L < dz) L = dz;
    double side2[OR_MAX_DEPTH + 1];
    uint64_t stride[OR_MAX_DEPTH + 1];
    for (int d = 0; d <= OR_MAX_DEPTH; ++d) {
        double side = L / (double)(1ULL << d);
        side2[d] = side * side;
        stride[d] = 1ULL << (63 - 3 * d);
    }
    int err = 0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
        uint64_t sp[256];
        uint8_t sd[256];
        int top = 0;
        int64_t na = 0, no = 0;
        double ax = 0.0, ay = 0.0, az = 0.0;
        const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        sp[top] = 0; sd[top] = 0; ++top;
        while (top) {
            --top;
            uint64_t pref = sp[top];
            int dep = sd[top];
            int64_t id = or_find(t, pref, dep);
            if (id < 0 || t->mass[id] == 0.0) continue;
            if (dep == OR_MAX_DEPTH && pref == keys[i]) continue;
            double ddx = t->cx[id] - px;
            double ddy = t->cy[id] - py;
            double ddz = t->cz[id] - pz;
            double r2 = ddx * ddx + ddy * ddy + ddz * ddz + soft2;
            if (dep == OR_MAX_DEPTH || side2[dep] < theta2 * r2) {
                double invR = 1.0 / sqrt(r2);
                double invR3 = invR * invR * invR;
                double s = G * t->mass[id] * invR3;
                ax += s * ddx; ay += s * ddy; az += s * ddz;
                ++na;
            } else {
                ++no;
                int cd = dep + 1;
                for (int oc = 0; oc < 8; ++oc) {
                    uint64_t child = pref + stride[cd] * (uint64_t)oc;
                    if (or_find(t, child, cd) >= 0) {
                        if (top < 256) { sp[top] = child; sd[top] = (uint8_t)cd; ++top; }
                        else { err = 1; }
                    }
                }
            }
        }
        acc[3 * i] = ax; acc[3 * i + 1] = ay; acc[3 * i + 2] = az;
        if (n_accept) n_accept[i] = na;
        if (n_open) n_open[i] = no;
    }
    return err ? -3 : OR_OK;
}

/* ------------------------------------------------------- direct sum --- */

/* FP64 direct-sum oracle: a_i = sum_{j != i, r_ij > 0} G m_j d / |d|^3 with
 * Neumaier-compensated sums in ascending j.  d = x_j - x_i; r^2 = |d|^2 +
 * soft2 formed as in traversal.cpp:72-82.  idx selects the sample bodies. */
int or_direct(int64_t n, const double *pos, const double *mass, double G, double soft2,
              int64_t m, const int64_t *idx, double *acc)
{
    if (n < 0 || m < 0 || (m > 0 && (!idx || !acc)) || (n > 0 && (!pos || !mass))) return OR_ERR_ARG;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t s = 0; s < m; ++s) {
        int64_t i = idx[s];
        double sum[3] = {0, 0, 0}, c[3] = {0, 0, 0};
        const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double d[3] = {pos[3 * j] - px, pos[3 * j + 1] - py, pos[3 * j + 2] - pz};
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + soft2;
            if (!(r2 > 0.0)) continue;
            double invR = 1.0 / sqrt(r2);
            double f = G * mass[j] * (invR * invR * invR);
            for (int a = 0; a < 3; ++a) {
                double term = f * d[a];
                double tt = sum[a] + term;
                if (fabs(sum[a]) >= fabs(term)) c[a] += (sum[a] - tt) + term;
                else c[a] += (term - tt) + sum[a];
                sum[a] = tt;
            }
        }
        for (int a = 0; a < 3; ++a) acc[3 * s + a] = sum[a] + c[a];
    }
    return OR_OK;
}

/* ------------------------------------------------------- step loop --- */

/* main.cpp:191-281 on one rank with --fc tree: drift-first KDK.  Runs
 * `nsteps` iterations of the loop body starting at time t0 (the first half
 * kick is skipped while t == 0, main.cpp:193).  State arrays are updated in
 * place and end in the last step's Morton order; acc holds the last forces.
 * Returns OR_OK or an error code. */
int or_simulate(int64_t n, double *pos, double *vel, double *mass, uint64_t *ids, double *acc,
                int64_t nsteps, double t0, double dt, double theta, double G, double soft2)
{
    if (n < 0 || nsteps < 0) return OR_ERR_ARG;
    if (n == 0) return OR_OK;
    uint64_t *keys = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    if (!keys) return OR_ERR_NOMEM;
    double t = t0;
    int rc = OR_OK;
    for (int64_t step = 0; step < nsteps; ++step, t += dt) {
        if (t > 0) {
            for (int64_t i = 0; i < 3 * n; ++i) vel[i] += 0.5 * dt * acc[i];
        }
        for (int64_t i = 0; i < 3 * n; ++i) pos[i] += dt * vel[i];
        double bb[6];
        or_bbox(n, pos, bb);
        or_morton_keys(n, pos, bb, keys);
        rc = or_sort(n, keys, pos, mass, vel, ids, NULL);
        if (rc) break;
        /* acc is indexed like the bodies; the reference recomputes it right
         * after the permutation, so its old order is never read */
        or_tree *tr = or_tree_build(n, keys, pos, mass);
        if (!tr) { rc = OR_ERR_NOMEM; break; }
        rc = or_walk(tr, n, keys, pos, bb, theta, G, soft2, acc, NULL, NULL);
        or_tree_free(tr);
        if (rc) break;
        for (int64_t i = 0; i < 3 * n; ++i) vel[i] += 0.5 * dt * acc[i];
    }
    free(keys);
    return rc;
}

/* main.cpp:191 with parse_time.h:6-16: the number of loop iterations that
 * `for (t = 0; t < t_end; t += dt)` performs. */
int64_t or_count_steps(double dt, double t_end)
{
    int64_t k = 0;
    for (double t = 0.0; t < t_end; t += dt) ++k;
    return k;
}
