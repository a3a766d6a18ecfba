/*
 * CPU oracle for the support-counting hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cupami_amd/) may load or
 * call this library; only tests/, the smoke check and bench.py's CPU-baseline
 * leg do, and only as the checker or as the timed CPU baseline.
 *
 * It restates the reference algorithms over a vertical store of 64-bit words
 * (row r of an item lives in word r/64, bit r%64; the reference uses 32-bit
 * words, which gives the same bit order and the same counts):
 *
 *   orc_bits_build      _bitsets.py:20-28, 91-100   (COO -> vertical store)
 *   orc_count_supports  _bitsets.py:30-53, 103-113  (complete intersection)
 *   orc_eclat           cuEclat.py:168-245          (prefix classes, retained rows)
 *   orc_apriori         clApriori.py:164-196, cuApriori.py:473-492,
 *                       _opencl_bitsets.py:21-42    (batches of 4096, u32 words)
 *   orc_periodic_eval   _temporal.py:28-55, 151-161 (support + max gap)
 *   orc_pfeclat         cuPFEclat.py:146-175, _temporal.py:163-168
 *
 * Patterns are returned as item column indexes (columns follow the string
 * order of labels, cuApriori.py:397), ascending within a pattern.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------ */
/* pattern lists                                                        */

typedef struct {
    uint32_t *items;     /* flattened item columns */
    uint64_t *offsets;   /* n+1 offsets into items */
    uint64_t *supports;
    int64_t  *periods;   /* NULL unless periodic */
    uint64_t n, cap_items, cap_n;
} orc_patterns;

static void pl_init(orc_patterns *p, int periodic) {
    memset(p, 0, sizeof(*p));
    p->cap_n = 64; p->cap_items = 256;
    p->items = malloc(p->cap_items * sizeof(uint32_t));
    p->offsets = malloc((p->cap_n + 1) * sizeof(uint64_t));
    p->supports = malloc(p->cap_n * sizeof(uint64_t));
    p->periods = periodic ? malloc(p->cap_n * sizeof(int64_t)) : NULL;
    p->offsets[0] = 0;
}

static void pl_push(orc_patterns *p, const uint32_t *prefix, uint32_t plen,
                    uint32_t last, uint64_t support, int64_t period) {
    if (p->n == p->cap_n) {
        p->cap_n *= 2;
        p->offsets = realloc(p->offsets, (p->cap_n + 1) * sizeof(uint64_t));
        p->supports = realloc(p->supports, p->cap_n * sizeof(uint64_t));
        if (p->periods) p->periods = realloc(p->periods, p->cap_n * sizeof(int64_t));
    }
    uint64_t need = p->offsets[p->n] + plen + 1;
    if (need > p->cap_items) {
        while (need > p->cap_items) p->cap_items *= 2;
        p->items = realloc(p->items, p->cap_items * sizeof(uint32_t));
    }
    uint32_t *dst = p->items + p->offsets[p->n];
    memcpy(dst, prefix, plen * sizeof(uint32_t));
    dst[plen] = last;
    p->supports[p->n] = support;
    if (p->periods) p->periods[p->n] = period;
    p->n++;
    p->offsets[p->n] = need;
}

static void pl_merge(orc_patterns *dst, orc_patterns *src) {
    for (uint64_t i = 0; i < src->n; ++i) {
        uint64_t a = src->offsets[i], b = src->offsets[i + 1];
        pl_push(dst, src->items + a, (uint32_t)(b - a - 1), src->items[b - 1],
                src->supports[i], src->periods ? src->periods[i] : 0);
    }
}

static void pl_free_fields(orc_patterns *p) {
    free(p->items); free(p->offsets); free(p->supports); free(p->periods);
}

uint64_t orc_patterns_count(const orc_patterns *p) { return p ? p->n : 0; }
uint64_t orc_patterns_items(const orc_patterns *p) { return p ? p->offsets[p->n] : 0; }

void orc_patterns_copy(const orc_patterns *p, uint32_t *items, uint64_t *offsets,
                       uint64_t *supports, int64_t *periods) {
    if (!p) return;
    memcpy(items, p->items, p->offsets[p->n] * sizeof(uint32_t));
    memcpy(offsets, p->offsets, (p->n + 1) * sizeof(uint64_t));
    memcpy(supports, p->supports, p->n * sizeof(uint64_t));
    if (periods && p->periods) memcpy(periods, p->periods, p->n * sizeof(int64_t));
}

void orc_patterns_free(orc_patterns *p) {
    if (!p) return;
    pl_free_fields(p);
    free(p);
}

/* ------------------------------------------------------------------ */
/* vertical store                                                       */

/* _bitsets.py:20-28: set bit (row) of item column; duplicates OR together. */
void orc_bits_build(const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                    uint64_t n_items, uint64_t words, uint64_t *bits) {
    memset(bits, 0, n_items * words * sizeof(uint64_t));
    for (uint64_t i = 0; i < nnz; ++i)
        bits[(uint64_t)cols[i] * words + rows[i] / 64] |= 1ull << (rows[i] % 64);
}

static uint64_t and_pop(const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t words) {
    uint64_t s = 0;
    for (uint64_t w = 0; w < words; ++w) {
        uint64_t v = a[w] & b[w];
        out[w] = v;
        s += (uint64_t)__builtin_popcountll(v);
    }
    return s;
}

static uint64_t pop_row(const uint64_t *a, uint64_t words) {
    uint64_t s = 0;
    for (uint64_t w = 0; w < words; ++w) s += (uint64_t)__builtin_popcountll(a[w]);
    return s;
}

/* _bitsets.py:30-53: support of each candidate = popcount of the AND of its rows. */
void orc_count_supports(const uint64_t *bits, uint64_t words, const uint32_t *cand,
                        uint64_t n, uint32_t width, uint64_t *out) {
    #pragma omp parallel for schedule(dynamic, 16)
    for (uint64_t c = 0; c < n; ++c) {
        const uint32_t *it = cand + c * width;
        uint64_t s = 0;
        for (uint64_t w = 0; w < words; ++w) {
            uint64_t v = ~0ull;
            for (uint32_t j = 0; j < width; ++j) v &= bits[(uint64_t)it[j] * words + w];
            s += (uint64_t)__builtin_popcountll(v);
        }
        out[c] = s;
    }
}

/* ------------------------------------------------------------------ */
/* Eclat (cuEclat.py:168-245)                                           */

typedef struct {
    uint64_t words, minsup;
    orc_patterns *out;
    uint32_t prefix[4096];
} eclat_ctx;

/* One prefix class: rows[k] are the retained tidsets of prefix+items[i]. */
static void eclat_class(eclat_ctx *cx, uint32_t plen, const uint32_t *items,
                        const uint64_t *rows, uint32_t k) {
    uint64_t W = cx->words;
    uint64_t *child = malloc((size_t)(k > 1 ? k - 1 : 1) * W * sizeof(uint64_t));
    uint32_t *citems = malloc((size_t)(k ? k : 1) * sizeof(uint32_t));
    for (uint32_t p = 0; p + 1 < k; ++p) {
        cx->prefix[plen] = items[p];
        uint32_t kept = 0;
        for (uint32_t q = p + 1; q < k; ++q) {
            uint64_t *dst = child + (uint64_t)kept * W;
            uint64_t s = and_pop(rows + (uint64_t)p * W, rows + (uint64_t)q * W, dst, W);
            if (s >= cx->minsup) {
                pl_push(cx->out, cx->prefix, plen + 1, items[q], s, 0);
                citems[kept++] = items[q];
            }
        }
        if (kept > 1) eclat_class(cx, plen + 1, citems, child, kept);
    }
    free(child);
    free(citems);
}

orc_patterns *orc_eclat(const uint64_t *bits, uint64_t words, uint32_t n_items, uint64_t minsup) {
    orc_patterns *res = malloc(sizeof(orc_patterns));
    pl_init(res, 0);
    uint32_t *freq = malloc((size_t)(n_items ? n_items : 1) * sizeof(uint32_t));
    uint32_t nf = 0;
    for (uint32_t i = 0; i < n_items; ++i) {
        uint64_t s = pop_row(bits + (uint64_t)i * words, words);
        if (s >= minsup) {
            pl_push(res, NULL, 0, i, s, 0);
            freq[nf++] = i;
        }
    }
    int nt = omp_get_max_threads();
    orc_patterns *parts = malloc(sizeof(orc_patterns) * (size_t)nt);
    for (int t = 0; t < nt; ++t) pl_init(&parts[t], 0);
    #pragma omp parallel
    {
        int t = omp_get_thread_num();
        eclat_ctx *cx = malloc(sizeof(eclat_ctx));
        cx->words = words; cx->minsup = minsup; cx->out = &parts[t];
        uint64_t *rows = malloc((size_t)(nf ? nf : 1) * words * sizeof(uint64_t));
        uint32_t *items = malloc((size_t)(nf ? nf : 1) * sizeof(uint32_t));
        #pragma omp for schedule(dynamic, 1)
        for (uint32_t a = 0; a < nf; ++a) {
            uint32_t left = freq[a], kept = 0;
            cx->prefix[0] = left;
            for (uint32_t b = a + 1; b < nf; ++b) {
                uint64_t s = and_pop(bits + (uint64_t)left * words,
                                     bits + (uint64_t)freq[b] * words,
                                     rows + (uint64_t)kept * words, words);
                if (s >= minsup) {
                    pl_push(cx->out, cx->prefix, 1, freq[b], s, 0);
                    items[kept++] = freq[b];
                }
            }
            if (kept > 1) eclat_class(cx, 1, items, rows, kept);
        }
        free(rows); free(items); free(cx);
    }
    for (int t = 0; t < nt; ++t) { pl_merge(res, &parts[t]); pl_free_fields(&parts[t]); }
    free(parts); free(freq);
    return res;
}

/* ------------------------------------------------------------------ */
/* Apriori, restating clApriori (the CPU-baseline algorithm)            */

static int cmp_pat(const uint32_t *a, const uint32_t *b, uint32_t k) {
    for (uint32_t i = 0; i < k; ++i) if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

/* binary search of a k-pattern in a sorted level */
static int level_has(const uint32_t *level, uint64_t n, uint32_t k, const uint32_t *pat) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        int c = cmp_pat(level + mid * k, pat, k);
        if (c == 0) return 1;
        if (c < 0) lo = mid + 1; else hi = mid;
    }
    return 0;
}

/* cuApriori.py:473-492: join sorted (k)-patterns sharing a (k-1)-prefix, keep a
 * candidate only if every k-subset is frequent.  Output is sorted because the
 * level is sorted and joins are emitted in (left, right) order. */
static uint32_t *gen_candidates(const uint32_t *level, uint64_t n, uint32_t k, uint64_t *out_n) {
    uint64_t cap = 1024, m = 0;
    uint32_t *out = malloc(cap * (k + 1) * sizeof(uint32_t));
    uint32_t *sub = malloc((k + 1) * sizeof(uint32_t));
    for (uint64_t a = 0; a < n; ++a) {
        for (uint64_t b = a + 1; b < n; ++b) {
            if (k > 1 && cmp_pat(level + a * k, level + b * k, k - 1) != 0) break;
            uint32_t cand[k + 1];
            memcpy(cand, level + a * k, k * sizeof(uint32_t));
            cand[k] = level[b * k + k - 1];
            int ok = 1;
            for (uint32_t drop = 0; drop + 2 < k + 1 && ok; ++drop) {
                uint32_t t = 0;
                for (uint32_t i = 0; i <= k; ++i) if (i != drop) sub[t++] = cand[i];
                ok = level_has(level, n, k, sub);
            }
            if (!ok) continue;
            if (m == cap) { cap *= 2; out = realloc(out, cap * (k + 1) * sizeof(uint32_t)); }
            memcpy(out + m * (k + 1), cand, (k + 1) * sizeof(uint32_t));
            ++m;
        }
    }
    free(sub);
    *out_n = m;
    return out;
}

/* _opencl_bitsets.py:21-42 counts over 32-bit words; bits32 is the same store
 * viewed as uint32 (little endian), words32 = 2 * words. */
static void count_u32(const uint32_t *bits32, uint64_t words32, const uint32_t *cand,
                      uint64_t n, uint32_t width, uint64_t *out) {
    #pragma omp parallel for schedule(dynamic, 4)
    for (uint64_t c = 0; c < n; ++c) {
        const uint32_t *it = cand + c * width;
        uint64_t s = 0;
        for (uint64_t w = 0; w < words32; ++w) {
            uint32_t v = 0xffffffffu;
            for (uint32_t j = 0; j < width; ++j) v &= bits32[(uint64_t)it[j] * words32 + w];
            s += (uint64_t)__builtin_popcount(v);
        }
        out[c] = s;
    }
}

/* clApriori.py:183-192: level loop, batches of batch (4096) candidates,
 * accept support > 0 and support >= minsup. */
orc_patterns *orc_apriori(const uint64_t *bits, uint64_t words, uint32_t n_items,
                          uint64_t minsup, uint64_t batch) {
    orc_patterns *res = malloc(sizeof(orc_patterns));
    pl_init(res, 0);
    if (batch == 0) batch = 4096;
    const uint32_t *b32 = (const uint32_t *)bits;
    uint64_t n = n_items;
    uint32_t k = 1;
    uint32_t *cand = malloc((n ? n : 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n_items; ++i) cand[i] = i;
    uint64_t *sup = malloc(batch * sizeof(uint64_t));
    while (n) {
        uint32_t *level = malloc(n * k * sizeof(uint32_t));
        uint64_t m = 0;
        for (uint64_t start = 0; start < n; start += batch) {
            uint64_t cnt = n - start < batch ? n - start : batch;
            count_u32(b32, 2 * words, cand + start * k, cnt, k, sup);
            for (uint64_t i = 0; i < cnt; ++i) {
                if (sup[i] > 0 && sup[i] >= minsup) {
                    const uint32_t *pat = cand + (start + i) * k;
                    pl_push(res, pat, k - 1, pat[k - 1], sup[i], 0);
                    memcpy(level + m * k, pat, k * sizeof(uint32_t));
                    ++m;
                }
            }
        }
        free(cand);
        cand = gen_candidates(level, m, k, &n);
        free(level);
        ++k;
    }
    free(cand); free(sup);
    return res;
}

/* ------------------------------------------------------------------ */
/* periodicity (_temporal.py:28-55)                                     */

/* Walk the set bits of one AND-ed row in row order: support and the largest
 * gap of [0, ts..., end].  Rows at or past n are ignored (never set). */
static void periodic_row(const uint64_t *row, uint64_t words, const int64_t *ts, uint64_t n,
                         int64_t end, uint64_t *sup_out, int64_t *per_out) {
    uint64_t s = 0;
    int64_t prev = 0, per = 0;
    for (uint64_t w = 0; w < words; ++w) {
        uint64_t v = row[w];
        while (v) {
            uint64_t r = w * 64 + (uint64_t)__builtin_ctzll(v);
            if (r < n) {
                int64_t t = ts[r];
                if (t - prev > per) per = t - prev;
                prev = t;
                ++s;
            }
            v &= v - 1;
        }
    }
    if (end - prev > per) per = end - prev;
    *sup_out = s;
    *per_out = per;
}

void orc_periodic_eval(const uint64_t *bits, uint64_t words, const int64_t *ts, uint64_t n,
                       int64_t end, const uint32_t *cand, uint64_t n_cand, uint32_t width,
                       uint64_t *sup, int64_t *per) {
    #pragma omp parallel
    {
        uint64_t *row = malloc(words * sizeof(uint64_t));
        #pragma omp for schedule(dynamic, 4)
        for (uint64_t c = 0; c < n_cand; ++c) {
            const uint32_t *it = cand + c * width;
            for (uint64_t w = 0; w < words; ++w) {
                uint64_t v = ~0ull;
                for (uint32_t j = 0; j < width; ++j) v &= bits[(uint64_t)it[j] * words + w];
                row[w] = v;
            }
            periodic_row(row, words, ts, n, end, &sup[c], &per[c]);
        }
        free(row);
    }
}

/* ------------------------------------------------------------------ */
/* PF-Eclat (cuPFEclat.py:146-175): accept = support >= minsup, support > 0,
 * period <= maxper; both conditions are anti-monotone, so the depth-first
 * search over retained rows visits exactly the accepted itemsets.        */

typedef struct {
    uint64_t words, minsup, n;
    int64_t maxper, end;
    const int64_t *ts;
    orc_patterns *out;
    uint32_t prefix[4096];
} pf_ctx;

static int pf_accept(pf_ctx *cx, uint64_t s, int64_t p) {
    return s >= cx->minsup && s > 0 && p <= cx->maxper;
}

static void pf_class(pf_ctx *cx, uint32_t plen, const uint32_t *items, const uint64_t *rows, uint32_t k) {
    uint64_t W = cx->words;
    uint64_t *child = malloc((size_t)(k > 1 ? k - 1 : 1) * W * sizeof(uint64_t));
    uint32_t *citems = malloc((size_t)(k ? k : 1) * sizeof(uint32_t));
    for (uint32_t p = 0; p + 1 < k; ++p) {
        cx->prefix[plen] = items[p];
        uint32_t kept = 0;
        for (uint32_t q = p + 1; q < k; ++q) {
            uint64_t *dst = child + (uint64_t)kept * W;
            and_pop(rows + (uint64_t)p * W, rows + (uint64_t)q * W, dst, W);
            uint64_t s; int64_t per;
            periodic_row(dst, W, cx->ts, cx->n, cx->end, &s, &per);
            if (pf_accept(cx, s, per)) {
                pl_push(cx->out, cx->prefix, plen + 1, items[q], s, per);
                citems[kept++] = items[q];
            }
        }
        if (kept > 1) pf_class(cx, plen + 1, citems, child, kept);
    }
    free(child);
    free(citems);
}

orc_patterns *orc_pfeclat(const uint64_t *bits, uint64_t words, uint32_t n_items,
                          const int64_t *ts, uint64_t n, int64_t end,
                          uint64_t minsup, int64_t maxper) {
    orc_patterns *res = malloc(sizeof(orc_patterns));
    pl_init(res, 1);
    uint32_t *freq = malloc((size_t)(n_items ? n_items : 1) * sizeof(uint32_t));
    uint32_t nf = 0;
    pf_ctx base = {0};
    base.words = words; base.minsup = minsup; base.n = n; base.maxper = maxper;
    base.end = end; base.ts = ts;
    for (uint32_t i = 0; i < n_items; ++i) {
        uint64_t s; int64_t per;
        periodic_row(bits + (uint64_t)i * words, words, ts, n, end, &s, &per);
        if (pf_accept(&base, s, per)) {
            pl_push(res, NULL, 0, i, s, per);
            freq[nf++] = i;
        }
    }
    int nt = omp_get_max_threads();
    orc_patterns *parts = malloc(sizeof(orc_patterns) * (size_t)nt);
    for (int t = 0; t < nt; ++t) pl_init(&parts[t], 1);
    #pragma omp parallel
    {
        int t = omp_get_thread_num();
        pf_ctx *cx = malloc(sizeof(pf_ctx));
        *cx = base;
        cx->out = &parts[t];
        uint64_t *rows = malloc((size_t)(nf ? nf : 1) * words * sizeof(uint64_t));
        uint32_t *items = malloc((size_t)(nf ? nf : 1) * sizeof(uint32_t));
        #pragma omp for schedule(dynamic, 1)
        for (uint32_t a = 0; a < nf; ++a) {
            uint32_t left = freq[a], kept = 0;
            cx->prefix[0] = left;
            for (uint32_t b = a + 1; b < nf; ++b) {
                uint64_t *dst = rows + (uint64_t)kept * words;
                and_pop(bits + (uint64_t)left * words, bits + (uint64_t)freq[b] * words, dst, words);
                uint64_t s; int64_t per;
                periodic_row(dst, words, ts, n, end, &s, &per);
                if (pf_accept(cx, s, per)) {
                    pl_push(cx->out, cx->prefix, 1, freq[b], s, per);
                    items[kept++] = freq[b];
                }
            }
            if (kept > 1) pf_class(cx, 1, items, rows, kept);
        }
        free(rows); free(items); free(cx);
    }
    for (int t = 0; t < nt; ++t) { pl_merge(res, &parts[t]); pl_free_fields(&parts[t]); }
    free(parts); free(freq);
    return res;
}

/* thread count actually used by the parallel regions */
int orc_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
