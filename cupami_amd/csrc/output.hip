// Pattern output: the key strings of a result and its save file, rendered in
// bulk from the result's forest and the label strings.  Replaces the Python
// loops that build getPatterns' keys and save's lines one pattern at a time
// (cuApriori.py:537-552, _temporal.py:170-181).
//
// A result is a forest: pattern i = pattern parent[i] plus column item[i],
// parents before children.  Columns are numbered in the string order of their
// labels, so ordering label tuples (Python's sorted() over tuples of str)
// is ordering column tuples, with no string comparison.  A key joins the
// labels of its columns in ascending order with "\t"; when every path in the
// forest ascends (the Eclat / Apriori searches), key(i) = key(parent) + "\t" +
// label(item), rendered in one pass; otherwise each pattern's columns are
// gathered and sorted first.

#include "cpm_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>

namespace cpm {
namespace {

constexpr uint32_t kNoParent = 0xFFFFFFFFu;

struct Labels {
    const char *text;
    const uint64_t *off;
    uint32_t n;
    uint64_t len(uint32_t c) const { return off[c + 1] - off[c]; }
    const char *at(uint32_t c) const { return text + off[c]; }
};

int check_forest(const cpm_patterns *p, const Labels &lab) {
    for (uint64_t i = 0; i < p->n; ++i) {
        const uint32_t par = p->parent[i];
        if (p->item[i] >= lab.n) return fail(CPM_ERR_INVALID, "pattern %llu names column %u of %u labels",
                                              (unsigned long long)i, p->item[i], lab.n);
        if (par != kNoParent && par >= i) return fail(CPM_ERR_INVALID, "pattern %llu: parent does not precede it",
                                                      (unsigned long long)i);
    }
    return CPM_OK;
}

// Does every child's column exceed its parent's?
bool ascending(const cpm_patterns *p) {
    for (uint64_t i = 0; i < p->n; ++i) {
        const uint32_t par = p->parent[i];
        if (par != kNoParent && p->item[i] <= p->item[par]) return false;
    }
    return true;
}

// Columns of every pattern, ascending: cols[start[i] .. start[i+1]).
void gather_columns(const cpm_patterns *p, std::vector<uint64_t> &start, std::vector<uint32_t> &cols) {
    std::vector<uint32_t> depth(p->n);
    start.assign(p->n + 1, 0);
    for (uint64_t i = 0; i < p->n; ++i) {
        depth[i] = p->parent[i] == kNoParent ? 1 : depth[p->parent[i]] + 1;
        start[i + 1] = start[i] + depth[i];
    }
    cols.resize(start[p->n]);
    for (uint64_t i = 0; i < p->n; ++i) {
        uint64_t k = start[i + 1];
        for (uint32_t j = (uint32_t)i; j != kNoParent; j = p->parent[j]) cols[--k] = p->item[j];
        std::sort(cols.begin() + start[i], cols.begin() + start[i + 1]);
    }
}

// Where the keys go: key i has koff[i+1] - koff[i] bytes.  For a forest
// whose paths do not ascend, each pattern's sorted columns are kept too.
struct KeyLayout {
    bool asc = true;
    std::vector<uint64_t> koff, start;
    std::vector<uint32_t> cols;
};

void key_layout(const cpm_patterns *p, const Labels &lab, KeyLayout &k) {
    k.asc = ascending(p);
    k.koff.assign(p->n + 1, 0);
    if (k.asc) {
        for (uint64_t i = 0; i < p->n; ++i) {
            const uint32_t par = p->parent[i];
            const uint64_t len = lab.len(p->item[i]) + (par == kNoParent ? 0 : k.koff[par + 1] - k.koff[par] + 1);
            k.koff[i + 1] = k.koff[i] + len;
        }
        return;
    }
    gather_columns(p, k.start, k.cols);
    for (uint64_t i = 0; i < p->n; ++i) {
        uint64_t len = k.start[i + 1] - k.start[i] - 1;
        for (uint64_t j = k.start[i]; j < k.start[i + 1]; ++j) len += lab.len(k.cols[j]);
        k.koff[i + 1] = k.koff[i] + len;
    }
}

// Key i is written at dst + koff[i] + i * gap, followed (gap == 1, not after
// the last key) by `sep`.  On an ascending forest a key copies its parent's,
// already written, and appends one label.
void key_render(const cpm_patterns *p, const Labels &lab, const KeyLayout &k, char *dst, uint64_t gap, char sep) {
    for (uint64_t i = 0; i < p->n; ++i) {
        char *d = dst + k.koff[i] + i * gap;
        if (k.asc) {
            const uint32_t par = p->parent[i];
            if (par != kNoParent) {
                const uint64_t plen = k.koff[par + 1] - k.koff[par];
                memcpy(d, dst + k.koff[par] + (uint64_t)par * gap, plen);
                d += plen;
                *d++ = '\t';
            }
            memcpy(d, lab.at(p->item[i]), lab.len(p->item[i]));
            d += lab.len(p->item[i]);
        } else {
            for (uint64_t j = k.start[i]; j < k.start[i + 1]; ++j) {
                if (j > k.start[i]) *d++ = '\t';
                memcpy(d, lab.at(k.cols[j]), lab.len(k.cols[j]));
                d += lab.len(k.cols[j]);
            }
        }
        if (gap && i + 1 < p->n) *d = sep;
    }
}

// Pattern indexes in ascending order of their column tuples.
std::vector<uint64_t> tuple_order(const cpm_patterns *p) {
    std::vector<uint64_t> start;
    std::vector<uint32_t> cols;
    gather_columns(p, start, cols);
    std::vector<uint64_t> order(p->n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
        return std::lexicographical_compare(cols.begin() + start[a], cols.begin() + start[a + 1],
                                            cols.begin() + start[b], cols.begin() + start[b + 1]);
    });
    return order;
}

int labels_in(const char *text, const uint64_t *off, uint32_t n, Labels &lab) {
    if (n && (!text || !off)) return fail(CPM_ERR_INVALID, "labels are NULL");
    if (n && off[0] != 0) return fail(CPM_ERR_INVALID, "label offsets must start at 0");
    for (uint32_t c = 0; c < n; ++c)
        if (off[c + 1] < off[c]) return fail(CPM_ERR_INVALID, "label offsets decrease");
    lab = {text, off, n};
    return CPM_OK;
}

}  // namespace
}  // namespace cpm

using namespace cpm;

extern "C" {

int cpm_patterns_keys(const cpm_patterns *p, const char *labels, const uint64_t *label_off, uint32_t n_labels,
                      char sep, char *out, uint64_t cap, uint64_t *bytes) {
    clear_error();
    if (!p || !bytes) return fail(CPM_ERR_INVALID, "null argument");
    Labels lab{};
    CPM_CHECK(labels_in(labels, label_off, n_labels, lab));
    CPM_CHECK(check_forest(p, lab));
    KeyLayout k;
    try {
        key_layout(p, lab, k);
    } catch (const std::bad_alloc &) {
        return fail(CPM_ERR_NOMEM, "out of host memory laying out %llu keys", (unsigned long long)p->n);
    }
    const uint64_t need = k.koff[p->n] + (p->n ? p->n - 1 : 0);
    *bytes = need;
    if (!out) return CPM_OK;
    if (cap < need) return fail(CPM_ERR_INVALID, "key buffer holds %llu bytes, %llu needed",
                                (unsigned long long)cap, (unsigned long long)need);
    key_render(p, lab, k, out, 1, sep);
    return CPM_OK;
}

int cpm_patterns_save(const cpm_patterns *p, const char *labels, const uint64_t *label_off, uint32_t n_labels,
                      const char *path) {
    clear_error();
    if (!p || !path) return fail(CPM_ERR_INVALID, "null argument");
    Labels lab{};
    CPM_CHECK(labels_in(labels, label_off, n_labels, lab));
    CPM_CHECK(check_forest(p, lab));
    KeyLayout k;
    std::string text;
    std::vector<uint64_t> order;
    try {
        key_layout(p, lab, k);
        text.resize(k.koff[p->n]);
        key_render(p, lab, k, &text[0], 0, 0);
        order = tuple_order(p);
    } catch (const std::bad_alloc &) {
        return fail(CPM_ERR_NOMEM, "out of host memory rendering %llu patterns", (unsigned long long)p->n);
    }
    FILE *f = fopen(path, "wb");
    if (!f) return fail(CPM_ERR_RUNTIME, "cannot open %s for writing", path);
    std::string buf;
    buf.reserve(1 << 20);
    bool ok = true;
    char num[64];
    for (uint64_t i : order) {
        buf.append(text.data() + k.koff[i], k.koff[i + 1] - k.koff[i]);
        int k = snprintf(num, sizeof num, " : %llu", (unsigned long long)p->support[i]);
        buf.append(num, (size_t)k);
        if (p->periodic) {
            k = snprintf(num, sizeof num, " : %lld", (long long)p->period[i]);
            buf.append(num, (size_t)k);
        }
        buf.push_back('\n');
        if (buf.size() >= (1u << 20)) {
            ok = ok && fwrite(buf.data(), 1, buf.size(), f) == buf.size();
            buf.clear();
        }
    }
    if (!buf.empty()) ok = ok && fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(CPM_ERR_RUNTIME, "writing %s failed", path);
    return CPM_OK;
}

}  // extern "C"
