// FP-growth search over device FP-trees (cuFPGrowth._search,
// cuFPGrowth.py:187-213; _fptree.conditional_tree, _fptree.py:180-231).
//
// The first tree is built level by level (fptree_build, fptree.hip).  Every
// later conditional tree whose frequent items fit one 64-bit mask is built by
// one workgroup of k_cond_small, all header items of a group of trees in one
// launch with one read-back per launch:
//
//   supports : the item's nodes walk to the root, adding their counts into
//              their ancestors' items (LDS atomics)
//   ranks    : the frequent ancestors ordered by (conditional support desc,
//              item id asc), the order fptree_build gives them
//   paths    : each node's frequent ancestors as a mask over ranks; the basket
//              of the conditional base is the mask's bits in rank order
//   sort     : masks sorted by their bit-reversed value, so that baskets
//              sharing a leading run of ranks sit next to each other
//   nodes    : a basket adds nodes from the lowest bit where its mask differs
//              from the one before it; one lane per rank walks the sorted
//              baskets, linking each new node to its parent and adding the
//              basket weights into the node it is on
//
// which is the prefix tree the reference's level-wise build makes of the same
// baskets (same nodes, counts and header), numbered depth-first instead of by
// level.  A conditional tree is kept in a compact layout: u32 parents, items
// as ranks in the tree, u32 counts (a count never exceeds the transaction
// count, held below 2^32), node lists per rank, and the per-rank sums of node
// depths that bound the size of the next conditional trees.  Items too many
// for a mask, or more nodes of one item than a workgroup sorts, take the
// general projection (fptree_conditional).  Single-path trees are enumerated
// on the host from their items and supports: the node counts of a single path
// are its items' conditional supports.  Conditional trees of one launch are
// searched depth-first, as the reference's stack of frames does; the
// patterns are the reference's (the order in which they are listed differs).
#include "fptree.h"

#include <string.h>
#include <memory>

namespace cpm {

namespace {

constexpr uint32_t kPaths = 2048;      // nodes of one item a workgroup projects
constexpr uint32_t kSrcItems = 2048;   // items of a source tree held in LDS
constexpr uint32_t kBits = 64;         // items of a tree built by one workgroup
constexpr uint32_t kCondThreads = 256;
constexpr uint32_t kPer = kPaths / kCondThreads;
constexpr uint32_t kNoItem = 0xffffffffu;
constexpr uint32_t kMaxTasks = 4096;            // workgroups of one launch
constexpr uint64_t kChunkWords = 32ull << 20;   // u32 words of the trees one launch builds

enum : uint32_t { kNone = 0, kOne = 1, kTree = 2, kLarge = 3 };

// One conditional tree to build: the source tree in the compact layout and
// where the new one goes.  out holds cap parents, cap items, cap counts,
// cap node-list entries, then kBits item ids.
struct CondTask {
    const uint32_t *par, *lit, *cnt;  // source tree
    const uint32_t *nodes;            // the source nodes of the item, n_x of them
    const uint32_t *gid;              // item id of each source rank
    uint32_t *out;
    uint32_t n_x, n_src, cap, pad;
};

// What a workgroup reports.  kOne: gid[0], sup[0] are the one frequent item.
// kTree: n_items ranks, their ids and supports, link[r]..link[r+1] the node
// list of rank r, bound[r] the sum over its nodes of their ancestors below
// the root.
struct CondResult {
    uint32_t status, n_items, n_nodes, single;
    uint32_t gid[kBits], sup[kBits], bound[kBits], link[kBits + 4];
};

__global__ __launch_bounds__(kCondThreads) void k_cond_small(const CondTask *__restrict__ tasks, uint32_t minsup,
                                                             CondResult *__restrict__ results) {
    __shared__ uint64_t sk[kPaths];   // bit-reversed masks, sorted
    __shared__ uint32_t si[kPaths];
    __shared__ uint32_t wv[kPaths];   // basket weights (in sorted order after the sort)
    __shared__ uint32_t bs[kPaths];   // first new node of each sorted basket
    __shared__ uint8_t bb[kPaths];    // lowest bit where a basket differs from the one before (64: equal)
    __shared__ uint32_t csup[kSrcItems];
    __shared__ uint8_t rmap[kSrcItems];   // source rank -> rank here (0xff: infrequent)
    __shared__ uint32_t fitem[kBits], fsup[kBits], fgid[kBits];
    __shared__ uint32_t cntp[kBits], bnd[kBits], loff[kBits];
    __shared__ uint32_t lds[16];
    __shared__ uint32_t s_nf, s_maxlen;

    const CondTask t = tasks[blockIdx.x];
    CondResult *res = results + blockIdx.x;
    const uint32_t tid = threadIdx.x;
    auto report = [&](uint32_t status, uint32_t n_items) {
        if (tid == 0) {
            res->status = status;
            res->n_items = n_items;
            res->n_nodes = 0;
            res->single = 1;
        }
    };
    if (t.n_x == 0 || t.n_x > kPaths || t.n_src > kSrcItems) {
        report(t.n_x == 0 ? kNone : kLarge, 0);
        return;
    }
    for (uint32_t i = tid; i < t.n_src; i += kCondThreads) {
        csup[i] = 0;
        rmap[i] = 0xff;
    }
    if (tid < kBits) {
        cntp[tid] = 0;
        bnd[tid] = 0;
    }
    if (tid == 0) {
        s_nf = 0;
        s_maxlen = 0;
    }
    __syncthreads();
    // conditional supports (parents precede their children, so p < c ends every walk)
    for (uint32_t k = tid; k < t.n_x; k += kCondThreads) {
        const uint32_t n = t.nodes[k], w = t.cnt[n];
        for (uint32_t c = n, p = t.par[n]; p != 0 && p < c; c = p, p = t.par[p]) atomicAdd(&csup[t.lit[p]], w);
    }
    __syncthreads();
    for (uint32_t i = tid; i < t.n_src; i += kCondThreads)
        if (csup[i] >= minsup) {
            const uint32_t slot = atomicAdd(&s_nf, 1u);
            if (slot < kBits) fitem[slot] = i;
        }
    __syncthreads();
    const uint32_t nf = s_nf;
    if (nf == 0 || nf > kBits) {
        report(nf ? kLarge : kNone, nf);
        return;
    }
    if (tid < nf) {
        fsup[tid] = csup[fitem[tid]];
        fgid[tid] = t.gid[fitem[tid]];
    }
    __syncthreads();
    uint32_t *const opar = t.out, *const olit = t.out + t.cap, *const ocnt = t.out + 2 * t.cap;
    if (tid < nf) {
        const uint32_t s = fsup[tid], g = fgid[tid];
        uint32_t r = 0;
        for (uint32_t f = 0; f < nf; ++f) r += (fsup[f] > s || (fsup[f] == s && fgid[f] < g)) ? 1u : 0u;
        rmap[fitem[tid]] = (uint8_t)r;
        res->gid[r] = g;
        res->sup[r] = s;
        t.out[4 * t.cap + r] = g;
    }
    if (nf == 1) {
        report(kOne, 1);
        return;
    }
    __syncthreads();
    // baskets: frequent ancestors as rank masks, sorted by the reversed mask
    uint32_t P = 1;
    while (P < t.n_x) P <<= 1;
    for (uint32_t k = tid; k < P; k += kCondThreads) {
        if (k < t.n_x) {
            const uint32_t n = t.nodes[k];
            uint64_t mask = 0;
            for (uint32_t c = n, p = t.par[n]; p != 0 && p < c; c = p, p = t.par[p]) {
                const uint32_t r = rmap[t.lit[p]];
                if (r != 0xff) mask |= 1ull << r;
            }
            sk[k] = __brevll(mask);
            wv[k] = t.cnt[n];
        } else {
            sk[k] = ~0ull;
        }
        si[k] = k;
    }
    __syncthreads();
    bitonic_sort_lds(sk, si, P);
    // new nodes per basket: from the lowest bit where it differs from the one before
    const uint32_t per = (t.n_x + kCondThreads - 1) / kCondThreads;
    const uint32_t i0 = tid * per;
    uint32_t wr[kPer], br[kPer], nr[kPer];
    uint32_t mine = 0, longest = 0;
#pragma unroll
    for (uint32_t e = 0; e < kPer; ++e) {
        const uint32_t i = i0 + e;
        nr[e] = 0;
        if (e < per && i < t.n_x) {
            const uint64_t m = __brevll(sk[i]);
            uint32_t b = 0;
            if (i) {
                const uint64_t d = m ^ __brevll(sk[i - 1]);
                b = d ? (uint32_t)__builtin_ctzll(d) : 64u;
            }
            const uint32_t len = __popcll(m);
            const uint32_t lcp = b >= 64 ? len : __popcll(m & ((1ull << b) - 1));
            nr[e] = len - lcp;
            br[e] = b;
            wr[e] = wv[si[i]];
            mine += nr[e];
            longest = max(longest, len);
        }
    }
    atomicMax(&s_maxlen, longest);
    uint32_t total;
    uint32_t base = 1 + block_excl_scan_u32(mine, lds, &total);
    if (1ull + total > t.cap) {   // never with the caller's bound; reported rather than written past
        report(kLarge, nf);
        return;
    }
#pragma unroll
    for (uint32_t e = 0; e < kPer; ++e) {
        const uint32_t i = i0 + e;
        if (e < per && i < t.n_x) {
            bs[i] = base;
            bb[i] = (uint8_t)br[e];
            wv[i] = wr[e];
            base += nr[e];
            // nodes per rank and the sum of their depths below the root
            const uint64_t m = __brevll(sk[i]);
            uint64_t fresh = br[e] >= 64 ? 0 : m & ~((1ull << br[e]) - 1);
            while (fresh) {
                const uint32_t p = (uint32_t)__builtin_ctzll(fresh);
                fresh &= fresh - 1;
                atomicAdd(&cntp[p], 1u);
                atomicAdd(&bnd[p], (uint32_t)__popcll(m & ((1ull << p) - 1)));
            }
        }
    }
    __syncthreads();
    if (tid < kBits) {
        const uint32_t c = cntp[tid];
        uint32_t x = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (tid >= (uint32_t)d) x += y;
        }
        loff[tid] = x - c;
        if (tid < nf) {
            res->link[tid] = x - c;
            res->bound[tid] = bnd[tid];
        }
    }
    if (tid == 0) {
        res->status = kTree;
        res->n_items = nf;
        res->n_nodes = 1 + total;
        res->single = total == s_maxlen ? 1u : 0u;
        res->link[nf] = total;
        opar[0] = 0;
        olit[0] = kNoItem;
        ocnt[0] = 0;
    }
    __syncthreads();
    // wave 0, lane p = rank p: walk the sorted baskets in order
    if (tid < kBits) {
        const uint32_t p = tid;
        const uint64_t below = (1ull << p) - 1;
        uint32_t *const olink = t.out + 3 * t.cap + loff[p];
        uint32_t cur = 0, run = 0, wrote = 0;
        bool have = false;
        for (uint32_t i = 0; i < t.n_x; ++i) {
            const uint64_t m = __brevll(sk[i]);
            const uint32_t b = bb[i];
            const uint64_t lm = m & below;
            const int q = lm ? 63 - __builtin_clzll(lm) : (int)p;
            const uint32_t before = __shfl(cur, q, 64);   // node of the rank before p, on the previous basket
            if ((m >> p) & 1ull) {
                if (p >= b) {   // b <= p < 64: a new node
                    if (have) ocnt[cur] = run;
                    const uint32_t lcp = __popcll(m & ((1ull << b) - 1));
                    const uint32_t j = __popcll(lm);
                    const uint32_t node = bs[i] + (j - lcp);
                    opar[node] = j == lcp ? (lm ? before : 0u) : node - 1;
                    olit[node] = p;
                    olink[wrote++] = node;
                    cur = node;
                    run = wv[i];
                    have = true;
                } else {
                    run += wv[i];
                }
            }
        }
        if (have) ocnt[cur] = run;
    }
}

__global__ void k_unit_weights(uint64_t *__restrict__ w, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        w[i] = 1;
}

// The general layout (int64 parents and items, u64 counts) in the compact
// one; local[] maps an item id to its rank in the tree.
__global__ void k_compact_tree(const int64_t *__restrict__ parents, const int64_t *__restrict__ items,
                               const unsigned long long *__restrict__ counts, uint64_t n,
                               const uint32_t *__restrict__ local, uint32_t *__restrict__ par,
                               uint32_t *__restrict__ lit, uint32_t *__restrict__ cnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == 0) {
            par[0] = 0;
            lit[0] = kNoItem;
            cnt[0] = 0;
        } else {
            par[i] = (uint32_t)parents[i];
            lit[i] = local[items[i]];
            cnt[i] = (uint32_t)counts[i];
        }
    }
}

// Per rank, the sum over its nodes of their ancestors below the root.
__global__ void k_rank_depths(const uint32_t *__restrict__ par, const uint32_t *__restrict__ lit, uint64_t n,
                              unsigned long long *__restrict__ bound) {
    for (uint64_t i = 1 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        unsigned long long d = 0;
        for (uint32_t c = (uint32_t)i, p = par[i]; p != 0 && p < c; c = p, p = par[p]) ++d;
        if (d) atomicAdd(bound + lit[i], d);
    }
}

// Device buffers handed out per launch and taken back when their trees are done.
struct Pool {
    std::vector<DBuf *> spare;
    ~Pool() {
        for (DBuf *b : spare) delete b;
    }
    std::shared_ptr<DBuf> get(size_t bytes, int *status) {
        size_t pick = spare.size();
        for (size_t i = 0; i < spare.size(); ++i)
            if (spare[i]->cap >= bytes && (pick == spare.size() || spare[i]->cap < spare[pick]->cap)) pick = i;
        DBuf *b;
        if (pick < spare.size()) {
            b = spare[pick];
            spare.erase(spare.begin() + pick);
        } else {
            b = new DBuf();
        }
        *status = b->ensure(std::max<size_t>(bytes, 256));
        if (*status != CPM_OK) {
            delete b;
            return nullptr;
        }
        return std::shared_ptr<DBuf>(b, [this](DBuf *x) { spare.push_back(x); });
    }
};

struct TreeDeleter {
    void operator()(cpm_fptree *t) const { cpm_fptree_destroy(t); }
};
using TreePtr = std::shared_ptr<cpm_fptree>;

// A tree still to be searched, in the compact layout.  Ranks run from the
// most to the least supported item; the header is visited from the last rank
// to the first (the reference's ascending-support header order).
struct Frame {
    std::shared_ptr<DBuf> mem;               // holds the arrays below
    const uint32_t *par = nullptr, *lit = nullptr, *cnt = nullptr, *gid = nullptr;
    const uint32_t *links = nullptr;         // node lists; rank r at links + lstart[r]
    uint64_t n_nodes = 0;
    uint32_t n_items = 0;
    std::vector<uint32_t> hgid, lstart, lcount;
    std::vector<uint64_t> hsup, bound;
    uint32_t pid = 0;                        // pattern id of the suffix
    uint32_t next = 0;                       // header positions scheduled so far
    TreePtr general;                         // the tree in the general layout, when there is one
};
using FramePtr = std::shared_ptr<Frame>;

struct Search {
    cpm_ctx *ctx;
    FpWs &ws;
    Pool pool;
    uint64_t minsup;
    cpm_patterns *res;
    std::vector<FramePtr> stack;
    DBuf d_tasks, d_results;
    HBuf h_tasks, h_results;
    static constexpr uint32_t kNoParent = 0xffffffffu;

    uint32_t emit(uint32_t parent, uint32_t item, uint64_t support) {
        res->parent.push_back(parent);
        res->item.push_back(item);
        res->support.push_back(support);
        return (uint32_t)(res->parent.size() - 1);
    }

    // every non-empty subset of a path of (item, count), support = the
    // smallest count; a subset's parent is the subset without its last item
    int enumerate_path(uint32_t pid, const uint32_t *items, const uint64_t *counts, uint64_t len) {
        if (len > 40) return fail(CPM_ERR_NOMEM, "single path of %llu items", (unsigned long long)len);
        struct Sub {
            uint64_t next;
            uint32_t pid;
            uint64_t minc;
        };
        std::vector<Sub> subs;
        subs.push_back({0, pid, ~0ull});
        while (!subs.empty()) {
            Sub &top = subs.back();
            if (top.next >= len) {
                subs.pop_back();
                continue;
            }
            const uint64_t p = top.next++;
            const uint64_t minc = std::min<uint64_t>(top.minc, counts[p]);
            const uint32_t id = emit(top.pid, items[p], minc);
            subs.push_back({p + 1, id, minc});
        }
        return CPM_OK;
    }

    // A tree in the general layout: its single path enumerated, or a frame
    // pushed onto `into`.
    int take_tree(TreePtr t, uint32_t pid, std::vector<FramePtr> &into) {
        hipStream_t st = ctx->stream;
        if (t->single_path) {
            std::vector<int64_t> it(t->n_nodes);
            std::vector<uint64_t> ct(t->n_nodes);
            CPM_CHECK(cpm_fptree_download(t.get(), nullptr, it.data(), ct.data(), nullptr, nullptr));
            std::vector<uint32_t> items(t->n_nodes);
            for (uint64_t i = 1; i < t->n_nodes; ++i) items[i] = (uint32_t)it[i];
            return enumerate_path(pid, items.data() + 1, ct.data() + 1, t->n_nodes - 1);
        }
        if (t->n_nodes > 0xffffffffull)
            return fail(CPM_ERR_NOMEM, "tree of %llu nodes", (unsigned long long)t->n_nodes);
        auto f = std::make_shared<Frame>();
        const uint32_t n = (uint32_t)t->header_item.size();
        f->general = t;
        f->pid = pid;
        f->n_nodes = t->n_nodes;
        f->n_items = n;
        f->hgid.resize(n);
        f->hsup.resize(n);
        std::vector<uint32_t> local(t->item_count, kNoItem);
        for (uint32_t r = 0; r < n; ++r) {   // header is ascending: rank r is its entry n - 1 - r
            f->hgid[r] = t->header_item[n - 1 - r];
            f->hsup[r] = t->header_sup[n - 1 - r];
            local[f->hgid[r]] = r;
        }
        CPM_CHECK(build_links(ctx, ws, t.get()));
        f->links = t->link_nodes.as<uint32_t>();
        f->lstart.resize(n);
        f->lcount.resize(n);
        for (uint32_t r = 0; r < n; ++r) {
            f->lstart[r] = t->link_off[f->hgid[r]];
            f->lcount[r] = t->link_off[f->hgid[r] + 1] - f->lstart[r];
        }
        // arrays: par, lit, cnt (n_nodes each), gid (n), local (item_count), bound (n u64)
        const uint64_t N = t->n_nodes;
        const uint64_t words = 3 * N + n + t->item_count;
        const size_t bound_off = (words * 4 + 7) / 8 * 8;
        int s = CPM_OK;
        f->mem = pool.get(bound_off + (size_t)std::max(n, 1u) * 8, &s);
        if (!f->mem) return s;
        uint32_t *w = f->mem->as<uint32_t>();
        uint32_t *par = w, *lit = w + N, *cnt = w + 2 * N, *gid = w + 3 * N, *loc = gid + n;
        unsigned long long *bound =
            reinterpret_cast<unsigned long long *>(f->mem->as<unsigned char>() + bound_off);
        CPM_HIP(hipMemcpyAsync(gid, f->hgid.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
        CPM_HIP(hipMemcpyAsync(loc, local.data(), (size_t)t->item_count * 4, hipMemcpyHostToDevice, st));
        CPM_HIP(hipMemsetAsync(bound, 0, (size_t)n * 8, st));
        hipLaunchKernelGGL(k_compact_tree, grid_of(N), dim3(256), 0, st, t->parents.as<int64_t>(),
                           t->items.as<int64_t>(), t->counts.as<unsigned long long>(), N, loc, par, lit, cnt);
        CPM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_rank_depths, grid_of(N), dim3(256), 0, st, par, lit, N, bound);
        CPM_HIP(hipGetLastError());
        f->bound.resize(n);
        CPM_CHECK(copy_to_host(ctx, f->bound.data(), bound, (size_t)n * 8));
        f->par = par;
        f->lit = lit;
        f->cnt = cnt;
        f->gid = gid;
        into.push_back(f);
        return CPM_OK;
    }

    // Conditional tree of rank x of f by the general projection.
    // Only trees in the general layout have items or nodes beyond what one
    // workgroup takes: a compact tree has at most kBits items and at most
    // kPaths nodes of any item (one per basket it was built from).
    int general_task(Frame &f, uint32_t x, uint32_t id, std::vector<FramePtr> &into) {
        if (!f.general) return fail(CPM_ERR_RUNTIME, "conditional tree of a compact tree exceeds its bound");
        cpm_fptree *cond = nullptr;
        CPM_CHECK(fptree_conditional(ctx, ws, f.general.get(), f.hgid[x], minsup, &cond));
        if (!cond) return CPM_OK;
        return take_tree(TreePtr(cond, TreeDeleter()), id, into);
    }

    struct Planned {
        FramePtr src;
        uint32_t x, id;
        uint64_t cap, offset;   // nodes, u32 words into the launch's buffer
        bool general;
    };

    // One launch over the header items of the frames at the top of the stack.
    int step() {
        std::vector<Planned> plan;
        uint64_t words = 0, launched = 0;
        while (!stack.empty() && launched < kMaxTasks && words < kChunkWords) {
            FramePtr f = stack.back();
            while (f->next < f->n_items && launched < kMaxTasks && words < kChunkWords) {
                const uint32_t x = f->n_items - 1 - f->next++;
                const uint32_t id = emit(f->pid, f->hgid[x], f->hsup[x]);
                if (f->lcount[x] == 0 || f->bound[x] == 0) continue;   // no ancestors: no conditional tree
                if (f->lcount[x] > kPaths || f->n_items > kSrcItems) {
                    plan.push_back({f, x, id, 0, 0, true});
                    continue;
                }
                const uint64_t cap = std::min<uint64_t>(f->bound[x], (uint64_t)f->lcount[x] * kBits) + 1;
                plan.push_back({f, x, id, cap, words, false});
                words += 4 * cap + kBits;
                ++launched;
            }
            if (f->next < f->n_items) break;
            stack.pop_back();
        }
        std::shared_ptr<DBuf> chunk;
        if (launched) {
            int s = CPM_OK;
            chunk = pool.get(words * 4, &s);
            if (!chunk) return s;
            CPM_CHECK(h_tasks.ensure(launched * sizeof(CondTask)));
            CPM_CHECK(d_tasks.ensure(launched * sizeof(CondTask)));
            CPM_CHECK(h_results.ensure(launched * sizeof(CondResult)));
            CPM_CHECK(d_results.ensure(launched * sizeof(CondResult)));
            CondTask *tk = h_tasks.as<CondTask>();
            uint64_t k = 0;
            for (const Planned &p : plan) {
                if (p.general) continue;
                const Frame &f = *p.src;
                tk[k++] = {f.par,       f.lit, f.cnt, f.links + f.lstart[p.x], f.gid, chunk->as<uint32_t>() + p.offset,
                           f.lcount[p.x], f.n_items, (uint32_t)p.cap, 0};
            }
            hipStream_t st = ctx->stream;
            CPM_HIP(hipMemcpyAsync(d_tasks.p, h_tasks.p, launched * sizeof(CondTask), hipMemcpyHostToDevice, st));
            const uint32_t ms = (uint32_t)std::min<uint64_t>(minsup, 0xffffffffull);
            hipLaunchKernelGGL(k_cond_small, dim3((unsigned)launched), dim3(kCondThreads), 0, st,
                               d_tasks.as<CondTask>(), ms, d_results.as<CondResult>());
            CPM_HIP(hipGetLastError());
            CPM_HIP(hipMemcpyAsync(h_results.p, d_results.p, launched * sizeof(CondResult), hipMemcpyDeviceToHost,
                                   st));
            CPM_HIP(hipStreamSynchronize(st));
        }
        // results in plan order; children are searched in that order, depth first
        std::vector<FramePtr> children;
        const CondResult *rs = h_results.as<CondResult>();
        uint64_t k = 0;
        for (const Planned &p : plan) {
            if (p.general) {
                CPM_CHECK(general_task(*p.src, p.x, p.id, children));
                continue;
            }
            const CondResult &r = rs[k++];
            if (r.status == kNone) continue;
            if (r.status == kOne) {
                emit(p.id, r.gid[0], r.sup[0]);
                continue;
            }
            if (r.status == kLarge) {
                CPM_CHECK(general_task(*p.src, p.x, p.id, children));
                continue;
            }
            const uint32_t n = r.n_items;
            if (r.single) {
                uint64_t sup[kBits];
                for (uint32_t i = 0; i < n; ++i) sup[i] = r.sup[i];
                CPM_CHECK(enumerate_path(p.id, r.gid, sup, n));
                continue;
            }
            auto f = std::make_shared<Frame>();
            uint32_t *out = chunk->as<uint32_t>() + p.offset;
            f->mem = chunk;
            f->par = out;
            f->lit = out + p.cap;
            f->cnt = out + 2 * p.cap;
            f->links = out + 3 * p.cap;
            f->gid = out + 4 * p.cap;
            f->n_nodes = r.n_nodes;
            f->n_items = n;
            f->pid = p.id;
            f->hgid.assign(r.gid, r.gid + n);
            f->hsup.assign(r.sup, r.sup + n);
            f->bound.assign(r.bound, r.bound + n);
            f->lstart.resize(n);
            f->lcount.resize(n);
            for (uint32_t i = 0; i < n; ++i) {
                f->lstart[i] = r.link[i];
                f->lcount[i] = r.link[i + 1] - r.link[i];
            }
            children.push_back(f);
        }
        for (size_t i = children.size(); i-- > 0;) stack.push_back(children[i]);
        return CPM_OK;
    }
};

}  // namespace

}  // namespace cpm

using namespace cpm;

extern "C" {

// FP-growth (cuFPGrowth._search, cuFPGrowth.py:170-213).
int cpm_fpgrowth_mine(cpm_ctx *ctx, const uint64_t *rows, const uint32_t *cols, uint64_t nnz,
                      uint64_t n_transactions, uint32_t n_items, uint64_t minsup, cpm_patterns **out) {
    clear_error();
    CPM_CHECK(bind(ctx));
    if (!out) return fail(CPM_ERR_INVALID, "null out");
    *out = nullptr;
    if (nnz && (!rows || !cols)) return fail(CPM_ERR_INVALID, "null argument");
    if (n_items && minsup < 1)
        return fail(CPM_ERR_INVALID, "FP-growth requires a positive minimum support count");
    if (n_transactions > 0xffffffffull)
        return fail(CPM_ERR_INVALID, "FP-growth takes fewer than 2^32 transactions");
    for (uint64_t i = 0; i < nnz; ++i)
        if (rows[i] >= n_transactions || cols[i] >= n_items)
            return fail(CPM_ERR_INVALID, "pair %llu outside %llu x %u", (unsigned long long)i,
                        (unsigned long long)n_transactions, n_items);
    hipStream_t st = ctx->stream;
    std::unique_ptr<cpm_patterns> res(new cpm_patterns());
    if (nnz == 0 || n_items == 0) {
        *out = res.release();
        return CPM_OK;
    }
    FpWs ws;
    DBuf dr, dc, ur, uc, w;
    CPM_CHECK(dr.ensure(nnz * 8));
    CPM_CHECK(dc.ensure(nnz * 4));
    CPM_CHECK(ur.ensure(nnz * 8));
    CPM_CHECK(uc.ensure(nnz * 4));
    CPM_CHECK(w.ensure(std::max<uint64_t>(n_transactions, 1) * 8));
    CPM_HIP(hipMemcpyAsync(dr.p, rows, nnz * 8, hipMemcpyHostToDevice, st));
    CPM_HIP(hipMemcpyAsync(dc.p, cols, nnz * 4, hipMemcpyHostToDevice, st));
    uint64_t nu = 0;
    CPM_CHECK(unique_pairs(ctx, ws, dr.as<uint64_t>(), dc.as<uint32_t>(), nnz, ur.as<uint64_t>(), uc.as<uint32_t>(),
                           &nu, n_transactions ? n_transactions - 1 : 0));
    hipLaunchKernelGGL(k_unit_weights, grid_of(n_transactions), dim3(256), 0, st, w.as<uint64_t>(),
                       n_transactions);
    CPM_HIP(hipGetLastError());
    cpm_fptree *root = nullptr;
    CPM_CHECK(fptree_build(ctx, ws, ur.as<uint64_t>(), uc.as<uint32_t>(), nu, w.as<uint64_t>(), n_transactions,
                           n_items, minsup, nullptr, &root));
    if (root) {
        Search search{ctx, ws, {}, minsup, res.get(), {}, {}, {}, {}, {}};
        CPM_CHECK(search.take_tree(TreePtr(root, TreeDeleter()), Search::kNoParent, search.stack));
        while (!search.stack.empty()) CPM_CHECK(search.step());
    }
    res->n = res->parent.size();
    *out = res.release();
    return CPM_OK;
}

}  // extern "C"
