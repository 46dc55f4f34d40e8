/*
 * mf_oracle.c -- CPU restatement of the two map-folding DFS formulations.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product (map_folding_amd/) never links, loads or
 * calls anything built from here.
 *
 * Restated from the reference (johnowhitaker/map_folding), file:line cited on
 * each function:
 *   Lunnon DFS ........ mf.c:14-163, mf_fast.cpp:53-259, gpu_folding_v2.cu:131-256
 *   symmetry DFS ...... sym_fold.cpp:122-527, experiments/9x9/sym_fold9.cpp:342-357
 *
 * Every routine carries exact event counters so that GPU kernels can be checked
 * not only on the count but also on the amount of tree they walked:
 *   Lunnon: inserts, gap computations, walk steps, candidates, completions
 *   sym:    nodes (= get_good calls), page walks, walk steps, leaves
 *
 * Pinning (which independent fact checks each function) is listed in
 * tests/test_oracle.py; no function here is "parity unpinned".
 *
 * Integer arithmetic only.  Totals that may exceed 64 bits are returned as
 * (lo, hi) pairs of uint64_t.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_LEAVES 64      /* n = prod(dims) <= 64 */
#define ORC_MAX_DIMS 8
#define ORC_MAX_GAP (ORC_MAX_LEAVES * ORC_MAX_LEAVES + 1)

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Lunnon DFS                                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t inserts;       /* leaf insertions (mf.c:136-142) */
    uint64_t gapcomps;      /* gap computations (mf.c:88-124) */
    uint64_t walk_steps;    /* iterations of the gap walk (mf.c:99-108) */
    uint64_t candidates;    /* distinct candidates pushed before the filter */
    uint64_t completions;   /* l > n events, each worth +n (mf.c:83-85,9-11) */
    uint64_t peak_gap;      /* high-water mark of the gap stack */
} orc_lunnon_stats;

typedef struct {
    int ndim;
    int n;
    int p[ORC_MAX_DIMS];
    /* d[i][l][m], i in [0,ndim), l,m in [0,n].  m = 0 is defined as 0
       (mf.c:52 leaves it uninitialised; gpu_folding_v2.cu:36 zero-fills). */
    int16_t d[ORC_MAX_DIMS][ORC_MAX_LEAVES + 1][ORC_MAX_LEAVES + 1];
} orc_lunnon_problem;

/* Tables: mf.c:29-75.  bigP[i] = prod p[<i]; c[i][m] = 1-based coordinate of
   leaf m in dimension i; d[i][l][m] = Lunnon's connection rule. */
static int orc_lunnon_setup(orc_lunnon_problem *pr, const int *dims, int ndim) {
    if (ndim < 1 || ndim > ORC_MAX_DIMS) return 1;
    long n = 1;
    for (int i = 0; i < ndim; i++) {
        if (dims[i] < 1) return 1;
        n *= dims[i];
        if (n > ORC_MAX_LEAVES) return 2;
    }
    memset(pr, 0, sizeof(*pr));
    pr->ndim = ndim;
    pr->n = (int)n;
    for (int i = 0; i < ndim; i++) pr->p[i] = dims[i];
    int bigP[ORC_MAX_DIMS + 1];
    bigP[0] = 1;
    for (int i = 1; i <= ndim; i++) bigP[i] = bigP[i - 1] * dims[i - 1];
    int c[ORC_MAX_DIMS][ORC_MAX_LEAVES + 1];
    for (int i = 0; i < ndim; i++)
        for (int m = 1; m <= pr->n; m++)
            c[i][m] = (m - 1) / bigP[i] - ((m - 1) / bigP[i + 1]) * dims[i] + 1;
    for (int i = 0; i < ndim; i++) {
        for (int l = 1; l <= pr->n; l++) {
            for (int m = 1; m <= l; m++) {
                int delta = c[i][l] - c[i][m];
                int v;
                if ((delta & 1) == 0)
                    v = (c[i][m] == 1) ? m : m - bigP[i];
                else
                    v = (c[i][m] == dims[i] || m + bigP[i] > l) ? m : m + bigP[i];
                pr->d[i][l][m] = (int16_t)v;
            }
        }
    }
    return 0;
}

/* Export of the d table for tests: out[i*(n+1)*(n+1) + l*(n+1) + m]. */
int orc_lunnon_table(const int *dims, int ndim, int16_t *out, int out_len) {
    orc_lunnon_problem pr;
    int rc = orc_lunnon_setup(&pr, dims, ndim);
    if (rc) return rc;
    int s = pr.n + 1;
    if (out_len < ndim * s * s) return 3;
    for (int i = 0; i < ndim; i++)
        for (int l = 0; l < s; l++)
            for (int m = 0; m < s; m++) out[(i * s + l) * s + m] = pr.d[i][l][m];
    return 0;
}

/* Gap computation for leaf l: mf.c:88-124.  Pushes the surviving gaps on
   gap[g0..) in walk order and returns the new stack top. */
static int orc_gaps(const orc_lunnon_problem *pr, const int *b, int *count,
                    int *gap, int l, int g0, orc_lunnon_stats *st) {
    int dd = 0;
    int gg = g0;
    for (int i = 0; i < pr->ndim; i++) {
        int m = pr->d[i][l][l];
        if (m == l) {
            dd++;
        } else {
            while (m != l) {
                st->walk_steps++;
                gap[gg] = m;
                if (count[m]++ == 0) gg++;
                m = pr->d[i][l][b[m]];
            }
        }
    }
    if (dd == pr->ndim)
        for (int m = 0; m < l; m++) gap[gg++] = m;
    st->gapcomps++;
    st->candidates += (uint64_t)(gg - g0);
    if ((uint64_t)gg > st->peak_gap) st->peak_gap = (uint64_t)gg;
    int g = g0;
    for (int j = g0; j < gg; j++) {
        gap[g] = gap[j];
        if (count[gap[j]] == pr->ndim - dd) g++;
        count[gap[j]] = 0;
    }
    return g;
}

/* Full count: the main loop of mf.c:78-143 with flag = 1, mod = 0.
   Returns the signed total of mf.c as an unsigned 128-bit value. */
static u128 orc_lunnon_full(const orc_lunnon_problem *pr, orc_lunnon_stats *st) {
    int n = pr->n;
    int a[ORC_MAX_LEAVES + 2] = {0}, b[ORC_MAX_LEAVES + 2] = {0};
    int count[ORC_MAX_LEAVES + 2] = {0}, gapter[ORC_MAX_LEAVES + 2] = {0};
    static __thread int gap[ORC_MAX_GAP];
    u128 total = 0;
    int g = 0, l = 1;
    while (l > 0) {
        if (l <= 1 || b[0] == 1) {
            if (l > n) {
                total += (u128)n;
                st->completions++;
            } else {
                g = orc_gaps(pr, b, count, gap, l, gapter[l - 1], st);
            }
        }
        while (l > 0 && g == gapter[l - 1]) {
            l--;
            b[a[l]] = b[l];
            a[b[l]] = a[l];
        }
        if (l > 0) {
            a[l] = gap[--g];
            b[l] = b[a[l]];
            b[a[l]] = l;
            a[b[l]] = l;
            gapter[l++] = g;
            st->inserts++;
        }
    }
    return total;
}

int orc_lunnon_count(const int *dims, int ndim, uint64_t *total_lohi, orc_lunnon_stats *st) {
    orc_lunnon_problem pr;
    int rc = orc_lunnon_setup(&pr, dims, ndim);
    if (rc) return rc;
    orc_lunnon_stats local;
    memset(&local, 0, sizeof(local));
    u128 t = orc_lunnon_full(&pr, &local);
    total_lohi[0] = (uint64_t)t;
    total_lohi[1] = (uint64_t)(t >> 64);
    if (st) *st = local;
    return 0;
}

/* Frontier collector in the canonical list encoding (mf_fast.cpp:185-221,
   gpu_folding_v2.cu:220-256): DFS to depth D; every node at l = D+1 that
   passes the normal-form filter is one state, emitted in DFS order.  A state
   is written as b[0..D] (the "below" links of the circular list over leaves
   0..D); a[] is its inverse.  out_b holds max_states rows of (D+1) bytes.
   Returns the number of states found (which may exceed max_states; only the
   first max_states are written).  `st` receives the counters of levels <= D. */
int64_t orc_lunnon_frontier(const int *dims, int ndim, int depth, uint8_t *out_b,
                            int64_t max_states, orc_lunnon_stats *st) {
    orc_lunnon_problem pr;
    if (orc_lunnon_setup(&pr, dims, ndim)) return -1;
    int n = pr.n;
    if (depth < 1 || depth > n) return -2;
    orc_lunnon_stats local;
    memset(&local, 0, sizeof(local));
    int a[ORC_MAX_LEAVES + 2] = {0}, b[ORC_MAX_LEAVES + 2] = {0};
    int count[ORC_MAX_LEAVES + 2] = {0}, gapter[ORC_MAX_LEAVES + 2] = {0};
    static __thread int gap[ORC_MAX_GAP];
    int64_t states = 0;
    int g = 0, l = 1;
    for (;;) {
        if (l <= 1 || b[0] == 1) {
            if (l > depth) {
                if (out_b && states < max_states)
                    for (int k = 0; k <= depth; k++) out_b[states * (depth + 1) + k] = (uint8_t)b[k];
                states++;
                g = gapter[l - 1];
            } else {
                g = orc_gaps(&pr, b, count, gap, l, gapter[l - 1], &local);
            }
        }
        while (g == gapter[l - 1]) {
            --l;
            if (l == 0) {
                if (st) *st = local;
                return states;
            }
            b[a[l]] = b[l];
            a[b[l]] = a[l];
        }
        a[l] = gap[--g];
        b[l] = b[a[l]];
        b[a[l]] = l;
        a[b[l]] = l;
        gapter[l] = g;
        ++l;
        local.inserts++;
    }
}

/* Resume one state: solve_from_state (mf_fast.cpp:142-183) with start_leaf =
   D+1, i.e. count every completion below the state and stop when the DFS
   returns to the state's own level.  Returns the subtree count (sum of +n). */
static u128 orc_lunnon_resume_one(const orc_lunnon_problem *pr, int depth, const uint8_t *sb,
                                  orc_lunnon_stats *st) {
    int n = pr->n;
    int a[ORC_MAX_LEAVES + 2] = {0}, b[ORC_MAX_LEAVES + 2] = {0};
    int count[ORC_MAX_LEAVES + 2] = {0}, gapter[ORC_MAX_LEAVES + 2] = {0};
    static __thread int gap[ORC_MAX_GAP];
    for (int k = 0; k <= depth; k++) b[k] = sb[k];
    for (int k = 0; k <= depth; k++) a[b[k]] = k;
    int start = depth + 1;
    int l = start;
    int g = 0;
    u128 total = 0;
    for (;;) {
        if (l <= 1 || b[0] == 1) {
            if (l > n) {
                total += (u128)n;
                st->completions++;
            } else {
                g = orc_gaps(pr, b, count, gap, l, gapter[l - 1], st);
            }
        }
        while (g == gapter[l - 1]) {
            if (l == start) return total;
            --l;
            b[a[l]] = b[l];
            a[b[l]] = a[l];
        }
        a[l] = gap[--g];
        b[l] = b[a[l]];
        b[a[l]] = l;
        a[b[l]] = l;
        gapter[l] = g;
        ++l;
        st->inserts++;
    }
}

/* Resume states [0, num) of a frontier; per-state count and node (insert)
   count into out_count / out_nodes (either may be NULL); summed counters into
   st.  Returns 0 or an error code. */
int orc_lunnon_resume(const int *dims, int ndim, int depth, const uint8_t *states_b, int64_t num,
                      uint64_t *out_count, uint64_t *out_nodes, uint64_t *total_lohi,
                      orc_lunnon_stats *st) {
    orc_lunnon_problem pr;
    int rc = orc_lunnon_setup(&pr, dims, ndim);
    if (rc) return rc;
    if (depth < 1 || depth > pr.n) return 4;
    orc_lunnon_stats sum;
    memset(&sum, 0, sizeof(sum));
    u128 total = 0;
    for (int64_t s = 0; s < num; s++) {
        orc_lunnon_stats one;
        memset(&one, 0, sizeof(one));
        u128 c = orc_lunnon_resume_one(&pr, depth, states_b + s * (depth + 1), &one);
        total += c;
        if (out_count) out_count[s] = (uint64_t)c;
        if (out_nodes) out_nodes[s] = one.inserts;
        sum.inserts += one.inserts;
        sum.gapcomps += one.gapcomps;
        sum.walk_steps += one.walk_steps;
        sum.candidates += one.candidates;
        sum.completions += one.completions;
        if (one.peak_gap > sum.peak_gap) sum.peak_gap = one.peak_gap;
    }
    if (total_lohi) {
        total_lohi[0] = (uint64_t)total;
        total_lohi[1] = (uint64_t)(total >> 64);
    }
    if (st) *st = sum;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Symmetry-reduced cycle DFS (sym_fold.cpp)                                */
/* ------------------------------------------------------------------------ */

#define SYM_MAX 64

typedef struct {
    uint64_t nodes;       /* get_good calls (sym_fold9.cpp:342-344) */
    uint64_t page_walks;  /* pages with edge[i] < i that were walked */
    uint64_t walk_steps;  /* iterations of the cycle walk (sym_fold.cpp:188-195) */
    uint64_t leaves;      /* sum of popcount(good) at i = n2-1 */
} orc_sym_stats;

typedef struct {
    int n, n2;
    uint8_t edge[4][SYM_MAX];
} orc_sym_problem;

/* Center-out spiral numbering: build_map, sym_fold.cpp:122-145.  map[x*h+y]. */
static void orc_build_map(int w, int h, int *map) {
    for (int k = 0; k < w * h; k++) map[k] = w * h;
    int x = 0, y = 0, dx = 1, dy = 0;
    for (int i = w * h - 1; i >= 0; --i) {
        map[x * h + y] = i;
        int x1 = x + dx, y1 = y + dy;
        if (x1 < 0 || x1 >= w || y1 < 0 || y1 >= h || map[x1 * h + y1] < w * h) {
            int t = dy;
            dy = dx;
            dx = -t;
            x += dx;
            y += dy;
        } else {
            x = x1;
            y = y1;
        }
    }
}

/* Four edge pages: make_edges, sym_fold.cpp:147-173.  Page x%2 holds the
   x-neighbour pairs, page 2 + y%2 the y-neighbour pairs; n2 = no neighbour. */
static void orc_sym_setup(orc_sym_problem *pr, int n) {
    pr->n = n;
    pr->n2 = n * n;
    memset(pr->edge, pr->n2, sizeof(pr->edge));
    int map[SYM_MAX];
    orc_build_map(n, n, map);
    for (int x = 0; x < n; x++)
        for (int y = 0; y < n; y++) {
            int idx = map[x * n + y];
            if (x > 0) {
                int o = map[(x - 1) * n + y];
                pr->edge[x % 2][idx] = (uint8_t)o;
                pr->edge[x % 2][o] = (uint8_t)idx;
            }
            if (y > 0) {
                int o = map[x * n + y - 1];
                pr->edge[2 + y % 2][idx] = (uint8_t)o;
                pr->edge[2 + y % 2][o] = (uint8_t)idx;
            }
        }
}

int orc_sym_edges(int n, uint8_t *out /* 4*64 */, int *map_out /* n*n, may be NULL */) {
    if (n < 1 || n > 8) return 1;
    orc_sym_problem pr;
    orc_sym_setup(&pr, n);
    memcpy(out, pr.edge, sizeof(pr.edge));
    if (map_out) orc_build_map(n, n, map_out);
    return 0;
}

static inline uint64_t orc_low_bits(int c) { return c == 64 ? ~0ULL : ((1ULL << c) - 1ULL); }

/* get_good, sym_fold.cpp:179-204: mask of legal insertion points j < i. */
static uint64_t orc_get_good(const orc_sym_problem *pr, const uint8_t *cycle, int i, orc_sym_stats *st) {
    uint64_t good = orc_low_bits(i);
    st->nodes++;
    for (int p = 0; p < 4; p++) {
        const uint8_t *edge = pr->edge[p];
        int other = edge[i];
        if (other >= i) continue;
        st->page_walks++;
        uint64_t allowed = 1ULL << other;
        int j = cycle[other];
        while (j != other) {
            st->walk_steps++;
            int j1 = edge[j];
            if (j1 < i) j = j1;
            allowed |= 1ULL << j;
            j = cycle[j];
        }
        good &= allowed;
        if (good == 0) return 0;
    }
    return good;
}

/* recurse_count, sym_fold.cpp:219-233 (children in ascending j via ctz,
   iter_good :206-217).  Counted with the node definition of sym_fold9. */
static uint64_t orc_recurse_count(const orc_sym_problem *pr, uint8_t *cycle, int i, orc_sym_stats *st) {
    uint64_t good = orc_get_good(pr, cycle, i, st);
    if (good == 0) return 0;
    if (i == pr->n2 - 1) {
        uint64_t c = (uint64_t)__builtin_popcountll(good);
        st->leaves += c;
        return c;
    }
    uint64_t out = 0;
    while (good) {
        int j = __builtin_ctzll(good);
        good &= good - 1;
        uint8_t next = cycle[j];
        cycle[i] = next;
        cycle[j] = (uint8_t)i;
        out += orc_recurse_count(pr, cycle, i + 1, st);
        cycle[j] = next;
    }
    return out;
}

/* Count one case: cycle[0..i-1] is a valid cycle over the first i cells. */
int orc_sym_count_case(int n, const uint8_t *cycle_in, int i, uint64_t *count, orc_sym_stats *st) {
    if (n < 3 || n > 8) return 1;
    orc_sym_problem pr;
    orc_sym_setup(&pr, n);
    if (i < 3 || i > pr.n2 - 1) return 2;
    uint8_t cycle[SYM_MAX];
    memset(cycle, 0, sizeof(cycle));
    memcpy(cycle, cycle_in, (size_t)i);
    orc_sym_stats local;
    memset(&local, 0, sizeof(local));
    *count = orc_recurse_count(&pr, cycle, i, &local);
    if (st) *st = local;
    return 0;
}

/* ---- center enumeration + D4 / reversal dedup (sym_fold.cpp:252-408) ---- */

typedef struct {
    uint8_t cyc[SYM_MAX];
} orc_cycle;

typedef struct {
    orc_cycle *v;
    int64_t len, cap;
} orc_cycle_vec;

static void orc_vec_push(orc_cycle_vec *vec, const uint8_t *c) {
    if (vec->len == vec->cap) {
        vec->cap = vec->cap ? vec->cap * 2 : 1024;
        vec->v = (orc_cycle *)realloc(vec->v, (size_t)vec->cap * sizeof(orc_cycle));
    }
    memcpy(vec->v[vec->len++].cyc, c, SYM_MAX);
}

/* recurse_generate, sym_fold.cpp:252-265. */
static void orc_generate(const orc_sym_problem *pr, uint8_t *cycle, int i, int stop, orc_cycle_vec *out,
                         orc_sym_stats *st) {
    if (i == stop) {
        orc_vec_push(out, cycle);
        return;
    }
    uint64_t good = orc_get_good(pr, cycle, i, st);
    while (good) {
        int j = __builtin_ctzll(good);
        good &= good - 1;
        uint8_t next = cycle[j];
        cycle[i] = next;
        cycle[j] = (uint8_t)i;
        orc_generate(pr, cycle, i + 1, stop, out, st);
        cycle[j] = next;
    }
}

/* generate_permutations, sym_fold.cpp:316-350: the 8 D4 maps of the c x c
   center in the reference's fixed order. */
static void orc_perms(int c, uint8_t perm[8][16]) {
    int map[16];
    orc_build_map(c, c, map);
    int rx[16], ry[16];
    for (int x = 0; x < c; x++)
        for (int y = 0; y < c; y++) {
            rx[map[x * c + y]] = x;
            ry[map[x * c + y]] = y;
        }
    int d = c - 1;
    for (int t = 0; t < 8; t++) {
        memset(perm[t], 255, 16);
        for (int i = 0; i < c * c; i++) {
            int x = rx[i], y = ry[i], tx = 0, ty = 0;
            switch (t) {
            case 0: tx = x; ty = y; break;
            case 1: tx = y; ty = x; break;
            case 2: tx = x; ty = d - y; break;
            case 3: tx = y; ty = d - x; break;
            case 4: tx = d - x; ty = y; break;
            case 5: tx = d - y; ty = x; break;
            case 6: tx = d - x; ty = d - y; break;
            default: tx = d - y; ty = d - x; break;
            }
            perm[t][i] = (uint8_t)map[tx * c + ty];
        }
    }
}

/* A small open-addressing set of 16-byte keys (make_key, sym_fold.cpp:352-359). */
typedef struct {
    uint8_t (*k)[16];
    uint8_t *used;
    int64_t cap, len;
} orc_keyset;

static uint64_t orc_key_hash(const uint8_t *k) {
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < 16; i++) {
        h ^= k[i];
        h *= 1099511628211ULL;
    }
    return h;
}

static void orc_keyset_init(orc_keyset *s, int64_t cap) {
    s->cap = cap;
    s->len = 0;
    s->k = (uint8_t(*)[16])calloc((size_t)cap, 16);
    s->used = (uint8_t *)calloc((size_t)cap, 1);
}

static int orc_keyset_find(const orc_keyset *s, const uint8_t *k) {
    uint64_t h = orc_key_hash(k) % (uint64_t)s->cap;
    while (s->used[h]) {
        if (memcmp(s->k[h], k, 16) == 0) return 1;
        h = (h + 1) % (uint64_t)s->cap;
    }
    return 0;
}

static int orc_keyset_insert(orc_keyset *s, const uint8_t *k) {
    uint64_t h = orc_key_hash(k) % (uint64_t)s->cap;
    while (s->used[h]) {
        if (memcmp(s->k[h], k, 16) == 0) return 0;
        h = (h + 1) % (uint64_t)s->cap;
    }
    s->used[h] = 1;
    memcpy(s->k[h], k, 16);
    s->len++;
    return 1;
}

static void orc_make_key(const uint8_t *cycle, int c2, uint8_t *key) {
    memset(key, 255, 16);
    for (int i = 0; i < c2; i++) key[i] = cycle[i];
}

/* orbit_keys (sym_fold.cpp:361-388) into `orbit` (a fresh small set);
   returns the orbit size. */
static int orc_orbit(const uint8_t *cycle, int c, uint8_t perm[8][16], uint8_t keys[16][16]) {
    int c2 = c * c, nk = 0;
    for (int t = 0; t < 8; t++) {
        uint8_t mapped[SYM_MAX], reversed[SYM_MAX];
        memcpy(mapped, cycle, SYM_MAX);
        for (int i = 0; i < c2; i++) mapped[perm[t][i]] = perm[t][cycle[i]];
        memcpy(reversed, mapped, SYM_MAX);
        int i = 0;
        do {
            reversed[mapped[i]] = (uint8_t)i;
            i = mapped[i];
        } while (i != 0);
        uint8_t k1[16], k2[16];
        orc_make_key(mapped, c2, k1);
        orc_make_key(reversed, c2, k2);
        int dup1 = 0, dup2 = 0;
        for (int q = 0; q < nk; q++) {
            if (memcmp(keys[q], k1, 16) == 0) dup1 = 1;
        }
        if (!dup1) memcpy(keys[nk++], k1, 16);
        for (int q = 0; q < nk; q++) {
            if (memcmp(keys[q], k2, 16) == 0) dup2 = 1;
        }
        if (!dup2) memcpy(keys[nk++], k2, 16);
    }
    return nk;
}

/* Seeds for n >= 6: all center cycles at i = c^2 (c = 4 for even n, 3 for
   odd), deduplicated in generation order (deduplicate_cases, :390-408), each
   with its orbit size as multiplier.  Writes up to max_seeds rows of 64 bytes
   to out_cycles and multipliers to out_mult; returns the seed count, and the
   raw center-cycle count in *raw. */
int64_t orc_sym_seeds(int n, uint8_t *out_cycles, uint64_t *out_mult, int64_t max_seeds, int64_t *raw) {
    if (n < 6 || n > 8) return -1;
    orc_sym_problem pr;
    orc_sym_setup(&pr, n);
    int c = (n % 2 == 0) ? 4 : 3;
    int c2 = c * c;
    uint8_t cycle[SYM_MAX];
    memset(cycle, 0, sizeof(cycle));
    cycle[0] = 1;
    cycle[1] = 2;
    cycle[2] = 0;
    orc_cycle_vec centers = {0, 0, 0};
    orc_sym_stats st;
    memset(&st, 0, sizeof(st));
    orc_generate(&pr, cycle, 3, c2, &centers, &st);
    if (raw) *raw = centers.len;
    uint8_t perm[8][16];
    orc_perms(c, perm);
    orc_keyset seen;
    orc_keyset_init(&seen, 4 * centers.len * 16 + 64);
    int64_t nseeds = 0;
    for (int64_t k = 0; k < centers.len; k++) {
        uint8_t key[16];
        orc_make_key(centers.v[k].cyc, c2, key);
        if (orc_keyset_find(&seen, key)) continue;
        uint8_t keys[16][16];
        int nk = orc_orbit(centers.v[k].cyc, c, perm, keys);
        if (nseeds < max_seeds) {
            if (out_cycles) memcpy(out_cycles + nseeds * SYM_MAX, centers.v[k].cyc, SYM_MAX);
            if (out_mult) out_mult[nseeds] = (uint64_t)nk;
        }
        nseeds++;
        for (int q = 0; q < nk; q++) orc_keyset_insert(&seen, keys[q]);
    }
    free(centers.v);
    free(seen.k);
    free(seen.used);
    return nseeds;
}

/* Streaming expansion (recurse_stream_cases, sym_fold.cpp:491-527): DFS from
   each seed (in seed order) from seed_i to stop_i, children in ascending j.
   The global case index counts cases in that order; cases with index in
   [start, start+max) are written (64-byte cycle + multiplier).  Returns the
   number written; *total_cases receives the number of indices visited. */
typedef struct {
    const orc_sym_problem *pr;
    int stop_i;
    uint64_t start, end, index;
    uint8_t *out_cycles;
    uint64_t *out_mult;
    uint64_t mult;
    int64_t written;
    int count_all;
    orc_sym_stats st;
} orc_stream_ctx;

static int orc_stream(orc_stream_ctx *cx, uint8_t *cycle, int i) {
    if (i == cx->stop_i) {
        if (cx->index >= cx->start && cx->index < cx->end) {
            if (cx->out_cycles) memcpy(cx->out_cycles + cx->written * SYM_MAX, cycle, SYM_MAX);
            if (cx->out_mult) cx->out_mult[cx->written] = cx->mult;
            cx->written++;
        }
        ++cx->index;
        return !cx->count_all && cx->index >= cx->end;
    }
    uint64_t good = orc_get_good(cx->pr, cycle, i, &cx->st);
    while (good) {
        int j = __builtin_ctzll(good);
        good &= good - 1;
        uint8_t next = cycle[j];
        cycle[i] = next;
        cycle[j] = (uint8_t)i;
        int stop = orc_stream(cx, cycle, i + 1);
        cycle[j] = next;
        if (stop) return 1;
    }
    return 0;
}

int64_t orc_sym_cases(int n, int expanded, uint64_t start, uint64_t max_cases, uint8_t *out_cycles,
                      uint64_t *out_mult, uint64_t *total_cases) {
    if (n < 6 || n > 8) return -1;
    int c = (n % 2 == 0) ? 4 : 3;
    int c2 = c * c;
    if (expanded < c2 || expanded > n * n - 1) return -2;
    orc_sym_problem pr;
    orc_sym_setup(&pr, n);
    int64_t raw = 0;
    int64_t ns = orc_sym_seeds(n, NULL, NULL, 0, &raw);
    uint8_t *seeds = (uint8_t *)malloc((size_t)ns * SYM_MAX);
    uint64_t *mults = (uint64_t *)malloc((size_t)ns * sizeof(uint64_t));
    orc_sym_seeds(n, seeds, mults, ns, &raw);
    orc_stream_ctx cx;
    memset(&cx, 0, sizeof(cx));
    cx.pr = &pr;
    cx.stop_i = expanded;
    cx.start = start;
    cx.end = max_cases == 0 ? UINT64_MAX : start + max_cases;
    cx.out_cycles = out_cycles;
    cx.out_mult = out_mult;
    cx.count_all = (total_cases != NULL);
    for (int64_t s = 0; s < ns; s++) {
        uint8_t cycle[SYM_MAX];
        memcpy(cycle, seeds + s * SYM_MAX, SYM_MAX);
        cx.mult = mults[s];
        if (orc_stream(&cx, cycle, c2)) break;
    }
    if (total_cases) *total_cases = cx.index;
    free(seeds);
    free(mults);
    return cx.written;
}

/* ---- parallel case solver (solve_cases_parallel, sym_fold.cpp:410-489) ---- */

typedef struct {
    const orc_sym_problem *pr;
    const uint8_t *cycles;
    const uint64_t *mults;
    int64_t num;
    int i;
    int64_t *next;
    pthread_mutex_t *mu;
    uint64_t *out_count;
    uint64_t *out_nodes;
    u128 total;
    orc_sym_stats st;
} orc_worker;

static void *orc_worker_main(void *arg) {
    orc_worker *w = (orc_worker *)arg;
    for (;;) {
        pthread_mutex_lock(w->mu);
        int64_t k = (*w->next)++;
        pthread_mutex_unlock(w->mu);
        if (k >= w->num) break;
        uint8_t cycle[SYM_MAX];
        memcpy(cycle, w->cycles + k * SYM_MAX, SYM_MAX);
        orc_sym_stats one;
        memset(&one, 0, sizeof(one));
        uint64_t c = orc_recurse_count(w->pr, cycle, w->i, &one);
        w->total += (u128)c * (u128)(w->mults ? w->mults[k] : 1);
        if (w->out_count) w->out_count[k] = c;
        if (w->out_nodes) w->out_nodes[k] = one.nodes;
        w->st.nodes += one.nodes;
        w->st.page_walks += one.page_walks;
        w->st.walk_steps += one.walk_steps;
        w->st.leaves += one.leaves;
    }
    return NULL;
}

/* Count num cases (64-byte cycles at level i) on `threads` threads.
   total_lohi = sum mult*count (NOT yet multiplied by n2). */
int orc_sym_solve(int n, int i, const uint8_t *cycles, const uint64_t *mults, int64_t num, int threads,
                  uint64_t *out_count, uint64_t *out_nodes, uint64_t *total_lohi, orc_sym_stats *st) {
    if (n < 3 || n > 8) return 1;
    orc_sym_problem pr;
    orc_sym_setup(&pr, n);
    if (i < 3 || i > pr.n2 - 1) return 2;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    int64_t next = 0;
    pthread_mutex_t mu;
    pthread_mutex_init(&mu, NULL);
    orc_worker *ws = (orc_worker *)calloc((size_t)threads, sizeof(orc_worker));
    pthread_t *tids = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        ws[t].pr = &pr;
        ws[t].cycles = cycles;
        ws[t].mults = mults;
        ws[t].num = num;
        ws[t].i = i;
        ws[t].next = &next;
        ws[t].mu = &mu;
        ws[t].out_count = out_count;
        ws[t].out_nodes = out_nodes;
        pthread_create(&tids[t], NULL, orc_worker_main, &ws[t]);
    }
    u128 total = 0;
    orc_sym_stats sum;
    memset(&sum, 0, sizeof(sum));
    for (int t = 0; t < threads; t++) {
        pthread_join(tids[t], NULL);
        total += ws[t].total;
        sum.nodes += ws[t].st.nodes;
        sum.page_walks += ws[t].st.page_walks;
        sum.walk_steps += ws[t].st.walk_steps;
        sum.leaves += ws[t].st.leaves;
    }
    pthread_mutex_destroy(&mu);
    free(ws);
    free(tids);
    if (total_lohi) {
        total_lohi[0] = (uint64_t)total;
        total_lohi[1] = (uint64_t)(total >> 64);
    }
    if (st) *st = sum;
    return 0;
}

/* Whole-square answer as printed by sym_fold (main, sym_fold.cpp:649-774):
   n = 1 -> 1, n = 2 -> 8, n <= 5 -> 2*n2*recurse_count(initial, 3),
   n >= 6 -> n2 * sum over seeds (or cases at `expanded` for n >= 7) of
   mult * recurse_count.  Cases [start, start+max) of the stream only
   (max = 0: all).  Returns 0 or an error. */
int orc_sym_answer(int n, int expanded, uint64_t start, uint64_t max_cases, int threads, uint64_t *answer_lohi,
                   orc_sym_stats *st) {
    u128 ans = 0;
    orc_sym_stats local;
    memset(&local, 0, sizeof(local));
    if (n < 1 || n > 8) return 1;
    if (n == 1) {
        ans = 1;
    } else if (n == 2) {
        ans = 8;
    } else if (n <= 5) {
        uint8_t cycle[SYM_MAX] = {1, 2, 0};
        uint64_t c = 0;
        orc_sym_count_case(n, cycle, 3, &c, &local);
        ans = (u128)2 * (u128)(n * n) * (u128)c;
    } else {
        int c = (n % 2 == 0) ? 4 : 3;
        int stop = (n >= 7) ? expanded : c * c;
        uint64_t tot = 0;
        int64_t cnt = orc_sym_cases(n, stop, start, max_cases, NULL, NULL, &tot);
        (void)cnt;
        uint64_t avail = tot > start ? tot - start : 0;
        if (max_cases != 0 && max_cases < avail) avail = max_cases;
        uint8_t *cyc = (uint8_t *)malloc((size_t)(avail ? avail : 1) * SYM_MAX);
        uint64_t *mul = (uint64_t *)malloc((size_t)(avail ? avail : 1) * sizeof(uint64_t));
        int64_t got = orc_sym_cases(n, stop, start, avail, cyc, mul, NULL);
        uint64_t t2[2];
        int rc = orc_sym_solve(n, stop, cyc, mul, got, threads, NULL, NULL, t2, &local);
        free(cyc);
        free(mul);
        if (rc) return rc;
        u128 s = ((u128)t2[1] << 64) | t2[0];
        ans = s * (u128)(n * n);
    }
    answer_lohi[0] = (uint64_t)ans;
    answer_lohi[1] = (uint64_t)(ans >> 64);
    if (st) *st = local;
    return 0;
}

/* ---- sym_kernel.c semantics (mfah/wasm/sym_kernel.c:212-230) ------------- */

/* Paths from level i to level stop: recurse_count of sym_kernel.c, one node
   per call including the calls that reach stop. */
static uint64_t orc_count_to(const orc_sym_problem *pr, uint8_t *cycle, int i, int stop, uint64_t *nodes,
                             orc_sym_stats *st) {
    ++*nodes;
    if (i == stop) return 1;
    uint64_t good = orc_get_good(pr, cycle, i, st);
    uint64_t out = 0;
    while (good) {
        int j = __builtin_ctzll(good);
        good &= good - 1;
        uint8_t next = cycle[j];
        cycle[i] = next;
        cycle[j] = (uint8_t)i;
        out += orc_count_to(pr, cycle, i + 1, stop, nodes, st);
        cycle[j] = next;
    }
    return out;
}

int orc_sym_count_to(int n, const uint8_t *cycle_in, int start, int stop, uint64_t *count, uint64_t *nodes) {
    if (n < 2 || n > 8) return 1;
    orc_sym_problem pr;
    orc_sym_setup(&pr, n);
    if (start < 1 || start > stop || stop > pr.n2) return 2;
    uint8_t cycle[SYM_MAX];
    memset(cycle, 0, sizeof(cycle));
    memcpy(cycle, cycle_in, (size_t)start);
    orc_sym_stats st;
    memset(&st, 0, sizeof(st));
    *nodes = 0;
    *count = orc_count_to(&pr, cycle, start, stop, nodes, &st);
    return 0;
}

/* All cycles at `depth` grown from the initial cycle 0->1->2->0 at cell 3, in
   DFS order (the prefixes a volunteer work unit carries). */
int64_t orc_sym_prefixes(int n, int depth, uint8_t *out, int64_t max_out) {
    if (n < 3 || n > 8) return -1;
    orc_sym_problem pr;
    orc_sym_setup(&pr, n);
    if (depth < 3 || depth > pr.n2) return -2;
    uint8_t cycle[SYM_MAX];
    memset(cycle, 0, sizeof(cycle));
    cycle[0] = 1;
    cycle[1] = 2;
    cycle[2] = 0;
    orc_cycle_vec v = {0, 0, 0};
    orc_sym_stats st;
    memset(&st, 0, sizeof(st));
    orc_generate(&pr, cycle, 3, depth, &v, &st);
    int64_t k = v.len < max_out ? v.len : max_out;
    if (out)
        for (int64_t q = 0; q < k; q++) memcpy(out + q * SYM_MAX, v.v[q].cyc, SYM_MAX);
    int64_t total = v.len;
    free(v.v);
    return total;
}
