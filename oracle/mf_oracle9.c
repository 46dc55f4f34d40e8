/*
 * mf_oracle9.c -- CPU restatement of the 9x9-capable symmetry DFS
 * (experiments/9x9/sym_fold9.cpp): 128-bit insertion masks, up to 81 cells,
 * center size c in {3, 4, 5} (MF_CENTER_N; default 5 for n = 9).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the GPU path; see oracle.py).
 * Cycles are stored in 128-byte rows (cells >= n*n unused).
 *
 * Pinned by the reference binary built from its own source
 * (oracle/_ref/sym_fold9, tests/golden/make_9x9.sh): totals n <= 6, the 9x9
 * center and stream counts, and 9x9 chunk answers.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define C9_CELLS 128   /* row bytes; n*n <= 81 */
#define C9_KEY 32      /* center keys: c*c <= 25 cells */

typedef struct {
    uint64_t nodes, page_walks, walk_steps, leaves;
} orc9_stats;

typedef struct {
    int n, n2;
    uint8_t edge[4][C9_CELLS];
} orc9_problem;

/* build_spiral_map (sym_fold9.cpp:123-146, as sym_fold.cpp:122-145): numbers
 * count down from w*h-1 along a walk from corner (0,0), so the center is 0. */
static void o9_spiral_map(int w, int h, int *map) {
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

static int o9_iabs(int v) { return v < 0 ? -v : v; }

/* build_map (sym_fold9.cpp:160-229): the spiral, or for odd squares with
 * MF_ORDER = reverse | opposite | axes_first | corners_first the rings of
 * Chebyshev radius r = 0..w/2, each in spiral order, those beyond radius
 * (MF_LOCK_CENTER_N - 1) / 2 (default 3) reordered; cells are numbered in
 * that sequence.  Any other MF_ORDER value keeps every ring in spiral order. */
static void o9_build_map(int w, int h, int *map) {
    o9_spiral_map(w, h, map);
    const char *ot = getenv("MF_ORDER");
    if (ot == NULL || ot[0] == 0 || strcmp(ot, "spiral") == 0 || w != h || w % 2 == 0) return;
    const char *lt = getenv("MF_LOCK_CENTER_N");
    int lock_radius = ((lt ? atoi(lt) : 3) - 1) / 2;
    if (lock_radius < 0) lock_radius = 0;
    const int center = w / 2, n2 = w * h;
    int bx[C9_CELLS], by[C9_CELLS];
    for (int x = 0; x < w; x++)
        for (int y = 0; y < h; y++) {
            bx[map[x * h + y]] = x;
            by[map[x * h + y]] = y;
        }
    int ox[C9_CELLS], oy[C9_CELLS], no = 0;
    for (int r = 0; r <= center; r++) {
        int rx[C9_CELLS], ry[C9_CELLS], k = 0;
        for (int q = 0; q < n2; q++) {
            const int ax = o9_iabs(bx[q] - center), ay = o9_iabs(by[q] - center);
            if ((ax > ay ? ax : ay) == r) {
                rx[k] = bx[q];
                ry[k] = by[q];
                k++;
            }
        }
        int tx[C9_CELLS], ty[C9_CELLS], m = 0;
        if (r > lock_radius && strcmp(ot, "reverse") == 0) {
            for (int q = k - 1; q >= 0; q--) { tx[m] = rx[q]; ty[m] = ry[q]; m++; }
        } else if (r > lock_radius && strcmp(ot, "opposite") == 0) {
            const int half = k / 2;
            for (int q = 0; q < half; q++) {
                tx[m] = rx[q]; ty[m] = ry[q]; m++;
                if (q + half < k) { tx[m] = rx[q + half]; ty[m] = ry[q + half]; m++; }
            }
            if (k % 2) { tx[m] = rx[k - 1]; ty[m] = ry[k - 1]; m++; }
        } else if (r > lock_radius && strcmp(ot, "axes_first") == 0) {
            /* stable by Manhattan distance: take distances in increasing order */
            for (int d = 0; d <= 2 * center; d++)
                for (int q = 0; q < k; q++)
                    if (o9_iabs(rx[q] - center) + o9_iabs(ry[q] - center) == d) { tx[m] = rx[q]; ty[m] = ry[q]; m++; }
        } else if (r > lock_radius && strcmp(ot, "corners_first") == 0) {
            for (int pass = 0; pass < 2; pass++)
                for (int q = 0; q < k; q++) {
                    const int corner = o9_iabs(rx[q] - center) == r && o9_iabs(ry[q] - center) == r;
                    if (corner == (pass == 0)) { tx[m] = rx[q]; ty[m] = ry[q]; m++; }
                }
        } else {
            for (int q = 0; q < k; q++) { tx[m] = rx[q]; ty[m] = ry[q]; m++; }
        }
        for (int q = 0; q < m; q++) { ox[no] = tx[q]; oy[no] = ty[q]; no++; }
    }
    for (int q = 0; q < n2; q++) map[ox[q] * h + oy[q]] = q;
}

/* make_edges (as sym_fold.cpp:147-173): page x%2 / 2 + y%2, n*n = none. */
static void o9_setup(orc9_problem *pr, int n) {
    pr->n = n;
    pr->n2 = n * n;
    memset(pr->edge, pr->n2, sizeof(pr->edge));
    int map[C9_CELLS];
    o9_build_map(n, n, map);
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

static inline u128 o9_low(int c) { return c >= 128 ? ~(u128)0 : (((u128)1 << c) - 1); }
static inline int o9_ctz(u128 v) {
    uint64_t lo = (uint64_t)v;
    return lo ? __builtin_ctzll(lo) : 64 + __builtin_ctzll((uint64_t)(v >> 64));
}
static inline int o9_pop(u128 v) { return __builtin_popcountll((uint64_t)v) + __builtin_popcountll((uint64_t)(v >> 64)); }

/* get_good (sym_fold9.cpp:286-311) with 128-bit masks. */
static u128 o9_good(const orc9_problem *pr, const uint8_t *cycle, int i, orc9_stats *st) {
    u128 good = o9_low(i);
    st->nodes++;
    for (int p = 0; p < 4; p++) {
        const uint8_t *edge = pr->edge[p];
        int other = edge[i];
        if (other >= i) continue;
        st->page_walks++;
        u128 allowed = (u128)1 << other;
        int j = cycle[other];
        while (j != other) {
            st->walk_steps++;
            int j1 = edge[j];
            if (j1 < i) j = j1;
            allowed |= (u128)1 << j;
            j = cycle[j];
        }
        good &= allowed;
        if (good == 0) return 0;
    }
    return good;
}

/* recurse_count to n*n (node = one get_good call). */
static u128 o9_count(const orc9_problem *pr, uint8_t *cycle, int i, orc9_stats *st) {
    u128 good = o9_good(pr, cycle, i, st);
    if (good == 0) return 0;
    if (i == pr->n2 - 1) {
        st->leaves += (uint64_t)o9_pop(good);
        return (u128)o9_pop(good);
    }
    u128 out = 0;
    while (good) {
        int j = o9_ctz(good);
        good &= good - 1;
        uint8_t next = cycle[j];
        cycle[i] = next;
        cycle[j] = (uint8_t)i;
        out += o9_count(pr, cycle, i + 1, st);
        cycle[j] = next;
    }
    return out;
}

int orc9_count_case(int n, const uint8_t *cycle_in, int level, uint64_t count[2], orc9_stats *st) {
    if (n < 3 || n > 9) return 1;
    orc9_problem pr;
    o9_setup(&pr, n);
    if (level < 3 || level > pr.n2 - 1) return 2;
    uint8_t cycle[C9_CELLS];
    memset(cycle, 0, sizeof(cycle));
    memcpy(cycle, cycle_in, (size_t)level);
    orc9_stats local;
    memset(&local, 0, sizeof(local));
    u128 c = o9_count(&pr, cycle, level, &local);
    count[0] = (uint64_t)c;
    count[1] = (uint64_t)(c >> 64);
    if (st) *st = local;
    return 0;
}

/* ---- center enumeration and dedup (sym_fold9.cpp choose_center_size and
   the center steps of main; generate_permutations / make_key / orbit_keys /
   deduplicate_cases as in sym_fold.cpp:316-408) ---- */

typedef struct {
    uint8_t (*v)[C9_CELLS];
    int64_t len, cap;
} o9_vec;

static void o9_push(o9_vec *vec, const uint8_t *c) {
    if (vec->len == vec->cap) {
        vec->cap = vec->cap ? vec->cap * 2 : 4096;
        vec->v = (uint8_t(*)[C9_CELLS])realloc(vec->v, (size_t)vec->cap * C9_CELLS);
    }
    memcpy(vec->v[vec->len++], c, C9_CELLS);
}

static void o9_generate(const orc9_problem *pr, uint8_t *cycle, int i, int stop, o9_vec *out, orc9_stats *st) {
    if (i == stop) {
        o9_push(out, cycle);
        return;
    }
    u128 good = o9_good(pr, cycle, i, st);
    while (good) {
        int j = o9_ctz(good);
        good &= good - 1;
        uint8_t next = cycle[j];
        cycle[i] = next;
        cycle[j] = (uint8_t)i;
        o9_generate(pr, cycle, i + 1, stop, out, st);
        cycle[j] = next;
    }
}

/* The 8 D4 maps of the c x c center on spiral numbers, reference order. */
static void o9_perms(int c, uint8_t perm[8][C9_KEY]) {
    int map[C9_KEY];
    o9_build_map(c, c, map);
    int rx[C9_KEY], ry[C9_KEY];
    for (int x = 0; x < c; x++)
        for (int y = 0; y < c; y++) {
            rx[map[x * c + y]] = x;
            ry[map[x * c + y]] = y;
        }
    int d = c - 1;
    for (int t = 0; t < 8; t++) {
        memset(perm[t], 255, C9_KEY);
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

typedef struct {
    uint8_t (*k)[C9_KEY];
    uint8_t *used;
    uint64_t cap;
} o9_set;

static uint64_t o9_hash(const uint8_t *k) {
    uint64_t h = 1469598103934665603ULL;
    for (int i = 0; i < C9_KEY; i++) {
        h ^= k[i];
        h *= 1099511628211ULL;
    }
    return h;
}

static int o9_find(const o9_set *s, const uint8_t *k) {
    uint64_t h = o9_hash(k) % s->cap;
    while (s->used[h]) {
        if (memcmp(s->k[h], k, C9_KEY) == 0) return 1;
        h = (h + 1) % s->cap;
    }
    return 0;
}

static void o9_insert(o9_set *s, const uint8_t *k) {
    uint64_t h = o9_hash(k) % s->cap;
    while (s->used[h]) {
        if (memcmp(s->k[h], k, C9_KEY) == 0) return;
        h = (h + 1) % s->cap;
    }
    s->used[h] = 1;
    memcpy(s->k[h], k, C9_KEY);
}

static void o9_key(const uint8_t *cycle, int c2, uint8_t *key) {
    memset(key, 255, C9_KEY);
    for (int i = 0; i < c2; i++) key[i] = cycle[i];
}

static int o9_orbit(const uint8_t *cycle, int c, uint8_t perm[8][C9_KEY], uint8_t keys[16][C9_KEY]) {
    int c2 = c * c, nk = 0;
    for (int t = 0; t < 8; t++) {
        uint8_t mapped[C9_CELLS], reversed[C9_CELLS];
        memcpy(mapped, cycle, C9_CELLS);
        for (int i = 0; i < c2; i++) mapped[perm[t][i]] = perm[t][cycle[i]];
        memcpy(reversed, mapped, C9_CELLS);
        int i = 0;
        do {
            reversed[mapped[i]] = (uint8_t)i;
            i = mapped[i];
        } while (i != 0);
        uint8_t k[2][C9_KEY];
        o9_key(mapped, c2, k[0]);
        o9_key(reversed, c2, k[1]);
        for (int w = 0; w < 2; w++) {
            int dup = 0;
            for (int q = 0; q < nk; q++)
                if (memcmp(keys[q], k[w], C9_KEY) == 0) dup = 1;
            if (!dup) memcpy(keys[nk++], k[w], C9_KEY);
        }
    }
    return nk;
}

int orc9_default_center(int n) { return n >= 9 ? 5 : ((n % 2 == 0) ? 4 : 3); }

static int o9_center_ok(int n, int c) { return c >= 3 && c <= n && c * c <= 25 && (c % 2) == (n % 2); }

/* Seeds: every center cycle at c*c, first representative of each orbit under
   the 8 maps and reversal, with the orbit size as multiplier; rows of 128 B. */
int64_t orc9_seeds(int n, int c, uint8_t *out, uint64_t *out_mult, int64_t max_seeds, int64_t *raw) {
    if (n < 6 || n > 9 || !o9_center_ok(n, c)) return -1;
    orc9_problem pr;
    o9_setup(&pr, n);
    int c2 = c * c;
    uint8_t cycle[C9_CELLS];
    memset(cycle, 0, sizeof(cycle));
    cycle[0] = 1;
    cycle[1] = 2;
    cycle[2] = 0;
    o9_vec centers = {0, 0, 0};
    orc9_stats st;
    memset(&st, 0, sizeof(st));
    o9_generate(&pr, cycle, 3, c2, &centers, &st);
    if (raw) *raw = centers.len;
    uint8_t perm[8][C9_KEY];
    o9_perms(c, perm);
    o9_set seen;
    seen.cap = (uint64_t)centers.len * 4 + 1024;
    seen.k = (uint8_t(*)[C9_KEY])calloc(seen.cap, C9_KEY);
    seen.used = (uint8_t *)calloc(seen.cap, 1);
    int64_t ns = 0;
    for (int64_t k = 0; k < centers.len; k++) {
        uint8_t key[C9_KEY];
        o9_key(centers.v[k], c2, key);
        if (o9_find(&seen, key)) continue;
        uint8_t keys[16][C9_KEY];
        int nk = o9_orbit(centers.v[k], c, perm, keys);
        if (ns < max_seeds) {
            if (out) memcpy(out + ns * C9_CELLS, centers.v[k], C9_CELLS);
            if (out_mult) out_mult[ns] = (uint64_t)nk;
        }
        ns++;
        for (int q = 0; q < nk; q++) o9_insert(&seen, keys[q]);
    }
    free(centers.v);
    free(seen.k);
    free(seen.used);
    return ns;
}

/* The stream of expanded cases (recurse_stream_cases): seeds in order, DFS to
   `expanded`, children ascending; window [start, start+max). */
typedef struct {
    const orc9_problem *pr;
    int stop;
    uint64_t start, end, index;
    uint8_t *out;
    uint64_t *out_mult, mult;
    int64_t written;
    orc9_stats st;
} o9_stream_ctx;

static int o9_stream(o9_stream_ctx *cx, uint8_t *cycle, int i) {
    if (i == cx->stop) {
        if (cx->index >= cx->start && cx->index < cx->end) {
            memcpy(cx->out + cx->written * C9_CELLS, cycle, C9_CELLS);
            cx->out_mult[cx->written] = cx->mult;
            cx->written++;
        }
        ++cx->index;
        return cx->index >= cx->end;
    }
    u128 good = o9_good(cx->pr, cycle, i, &cx->st);
    while (good) {
        int j = o9_ctz(good);
        good &= good - 1;
        uint8_t next = cycle[j];
        cycle[i] = next;
        cycle[j] = (uint8_t)i;
        int stop = o9_stream(cx, cycle, i + 1);
        cycle[j] = next;
        if (stop) return 1;
    }
    return 0;
}

int64_t orc9_cases(int n, int c, int expanded, uint64_t start, uint64_t max_cases, uint8_t *out,
                   uint64_t *out_mult) {
    if (n < 6 || n > 9 || !o9_center_ok(n, c)) return -1;
    if (expanded < c * c || expanded > n * n - 1) return -2;
    orc9_problem pr;
    o9_setup(&pr, n);
    int64_t raw = 0;
    int64_t ns = orc9_seeds(n, c, NULL, NULL, 0, &raw);
    uint8_t *seeds = (uint8_t *)malloc((size_t)ns * C9_CELLS);
    uint64_t *mults = (uint64_t *)malloc((size_t)ns * sizeof(uint64_t));
    orc9_seeds(n, c, seeds, mults, ns, &raw);
    o9_stream_ctx cx;
    memset(&cx, 0, sizeof(cx));
    cx.pr = &pr;
    cx.stop = expanded;
    cx.start = start;
    cx.end = start + max_cases;
    cx.out = out;
    cx.out_mult = out_mult;
    for (int64_t s = 0; s < ns; s++) {
        uint8_t cyc[C9_CELLS];
        memcpy(cyc, seeds + s * C9_CELLS, C9_CELLS);
        cx.mult = mults[s];
        if (o9_stream(&cx, cyc, c * c)) break;
    }
    free(seeds);
    free(mults);
    return cx.written;
}

/* Count cases on host threads: per-case counts (u64 each; a 9x9 case's count
   fits) and node counts, the sum of mult * count as 128 bits, and stats. */
typedef struct {
    const orc9_problem *pr;
    const uint8_t *cycles;
    const uint64_t *mults;
    int level;
    int64_t k;
    int64_t next;
    pthread_mutex_t mu;
    uint64_t *per_count, *per_nodes;
    u128 total;
    orc9_stats st;
} o9_job;

static void *o9_worker(void *arg) {
    o9_job *jb = (o9_job *)arg;
    for (;;) {
        pthread_mutex_lock(&jb->mu);
        int64_t s = jb->next++;
        pthread_mutex_unlock(&jb->mu);
        if (s >= jb->k) break;
        uint8_t cyc[C9_CELLS];
        memcpy(cyc, jb->cycles + s * C9_CELLS, C9_CELLS);
        orc9_stats st;
        memset(&st, 0, sizeof(st));
        u128 c = o9_count(jb->pr, cyc, jb->level, &st);
        if (jb->per_count) jb->per_count[s] = (uint64_t)c;
        if (jb->per_nodes) jb->per_nodes[s] = st.nodes;
        pthread_mutex_lock(&jb->mu);
        jb->total += c * (jb->mults ? jb->mults[s] : 1);
        jb->st.nodes += st.nodes;
        jb->st.page_walks += st.page_walks;
        jb->st.walk_steps += st.walk_steps;
        jb->st.leaves += st.leaves;
        pthread_mutex_unlock(&jb->mu);
    }
    return NULL;
}

int orc9_solve(int n, int level, const uint8_t *cycles, const uint64_t *mults, int64_t k, int threads,
               uint64_t *per_count, uint64_t *per_nodes, uint64_t total[2], orc9_stats *st) {
    if (n < 3 || n > 9) return 1;
    orc9_problem pr;
    o9_setup(&pr, n);
    if (level < 3 || level > pr.n2 - 1) return 2;
    o9_job jb;
    memset(&jb, 0, sizeof(jb));
    jb.pr = &pr;
    jb.cycles = cycles;
    jb.mults = mults;
    jb.level = level;
    jb.k = k;
    jb.per_count = per_count;
    jb.per_nodes = per_nodes;
    pthread_mutex_init(&jb.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, o9_worker, &jb);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&jb.mu);
    total[0] = (uint64_t)jb.total;
    total[1] = (uint64_t)(jb.total >> 64);
    if (st) *st = jb.st;
    return 0;
}
