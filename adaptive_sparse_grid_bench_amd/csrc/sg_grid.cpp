// Host-side grid construction, bijection and launch tables; see sg_grid.hpp.
#include "sg_grid.hpp"

#include <array>
#include <functional>

#include <algorithm>
#include <cstring>

namespace sg {

namespace {
thread_local std::string g_last_error;

uint64_t level_key(const int* l, int d) {
    uint64_t k = 0;
    for (int j = 0; j < d; ++j) k = k * 31u + static_cast<uint64_t>(l[j]) + 1u;
    return k;
}
uint64_t level_key(const uint8_t* l, int d) {
    uint64_t k = 0;
    for (int j = 0; j < d; ++j) k = k * 31u + static_cast<uint64_t>(l[j]) + 1u;
    return k;
}

// compact-order key: level sum, then l_{d-1}, l_{d-2}, ... ascending (asg.c:266-279)
bool compact_less(const uint8_t* a, const uint8_t* b, int d) {
    int sa = 0, sb = 0;
    for (int j = 0; j < d; ++j) { sa += a[j]; sb += b[j]; }
    if (sa != sb) return sa < sb;
    for (int j = d - 1; j >= 0; --j)
        if (a[j] != b[j]) return a[j] < b[j];
    return false;
}

void fill_offsets(Grid& g) {
    g.lidx.assign(g.n + 1, 0);
    g.cnt.assign(g.n + 1, 0);
    g.sub_base.resize(g.S);
    int64_t off = 0;
    int64_t q = 0;
    for (int s = 0; s < g.n; ++s) {
        g.lidx[s] = off;
        g.cnt[s] = q;
        while (q < g.S) {
            int sum = 0;
            for (int j = 0; j < g.d; ++j) sum += g.sub_levels[q * g.d + j];
            if (sum != s) break;
            g.sub_base[q] = static_cast<int32_t>(off);
            off += int64_t(1) << s;
            ++q;
        }
    }
    g.lidx[g.n] = off;
    g.cnt[g.n] = q;
    g.N = off;
    for (int j = 0; j < g.d; ++j) g.max_level[j] = 0;
    for (int64_t i = 0; i < g.S; ++i)
        for (int j = 0; j < g.d; ++j)
            g.max_level[j] = std::max<int>(g.max_level[j], g.sub_levels[i * g.d + j]);
}
}  // namespace

const char* status_name(int st) {
    switch (st) {
        case kOk: return "ok";
        case kInvalid: return "invalid argument";
        case kNoMem: return "out of memory";
        case kHip: return "HIP runtime error";
        case kRange: return "out of range";
        case kUnsupported: return "unsupported shape";
        default: return "unknown status";
    }
}
void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

// dp[i][s] = number of level vectors with i+1 components, sum s, l_j <= limits[j]
// (asg.c:108-129, in the direct form of asg.cu:79-89; the s > limits[0] entries of
// row 0 are 0 where asg.c leaves them unset).
int build_regular(Grid& g, int d, int n, const int* limits) {
    if (d < 1 || d > kMaxDims || n < 1 || n > kMaxLevelSum + 1) {
        set_error("build_regular: need 1 <= d <= 20 and 1 <= n <= 27");
        return kInvalid;
    }
    g = Grid();
    g.d = d;
    g.n = n;
    for (int j = 0; j < d; ++j) {
        int lim = limits ? limits[j] : n - 1;
        if (lim < 0) {
            set_error("build_regular: negative limit");
            return kInvalid;
        }
        g.limits[j] = std::min(lim, n - 1);
        if (g.limits[j] > kMaxLevel) {
            set_error("build_regular: a per-dimension level above 15 is not supported");
            return kUnsupported;
        }
    }
    g.dp.assign(size_t(d) * n, 0);
    for (int s = 0; s < n; ++s) g.dp[s] = (s <= g.limits[0]) ? 1 : 0;
    for (int i = 1; i < d; ++i)
        for (int s = 0; s < n; ++s) {
            int64_t acc = 0;
            for (int k = 0; k <= s && k <= g.limits[i]; ++k) acc += g.dp[(i - 1) * n + (s - k)];
            g.dp[size_t(i) * n + s] = acc;
        }
    int64_t S = 0, N = 0;
    for (int s = 0; s < n; ++s) {
        S += g.dp[size_t(d - 1) * n + s];
        N += g.dp[size_t(d - 1) * n + s] << s;
    }
    if (N >= (int64_t(1) << 31)) {
        set_error("build_regular: grid has 2^31 points or more (offsets are int32)");
        return kRange;
    }
    g.S = S;
    g.sub_levels.resize(size_t(S) * d);
    int64_t q = 0;
    int lv[kMaxDims];
    for (int s = 0; s < n; ++s)
        for (int64_t r = 0; r < g.dp[size_t(d - 1) * n + s]; ++r, ++q) {
            idx2l(g, r, s, lv);
            for (int j = 0; j < d; ++j) g.sub_levels[q * d + j] = static_cast<uint8_t>(lv[j]);
        }
    fill_offsets(g);
    return kOk;
}

int build_explicit(Grid& g, int d, const uint8_t* levels, int64_t count) {
    if (d < 1 || d > kMaxDims || count < 1) {
        set_error("build_explicit: need 1 <= d <= 20 and a non-empty level set");
        return kInvalid;
    }
    g = Grid();
    g.d = d;
    g.explicit_set = true;
    std::vector<int64_t> order(count);
    for (int64_t i = 0; i < count; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return compact_less(levels + a * d, levels + b * d, d);
    });
    int maxsum = 0;
    g.sub_levels.clear();
    for (int64_t i = 0; i < count; ++i) {
        const uint8_t* l = levels + order[i] * d;
        if (i > 0 && std::memcmp(l, levels + order[i - 1] * d, d) == 0) continue;  // dedupe
        int sum = 0;
        for (int j = 0; j < d; ++j) {
            if (l[j] > kMaxLevel) {
                set_error("build_explicit: a per-dimension level above 15 is not supported");
                return kUnsupported;
            }
            sum += l[j];
        }
        maxsum = std::max(maxsum, sum);
        g.sub_levels.insert(g.sub_levels.end(), l, l + d);
    }
    if (maxsum > kMaxLevelSum) {
        set_error("build_explicit: level sum above 26 is not supported");
        return kUnsupported;
    }
    g.S = int64_t(g.sub_levels.size()) / d;
    g.n = maxsum + 1;
    for (int64_t i = 0; i < g.S; ++i) g.index_of[level_key(&g.sub_levels[i * d], d)] = i;
    // downward closed: every l - e_k (l_k > 0) must be present
    int lv[kMaxDims];
    for (int64_t i = 0; i < g.S; ++i) {
        for (int j = 0; j < d; ++j) lv[j] = g.sub_levels[i * d + j];
        for (int j = 0; j < d; ++j) {
            if (lv[j] == 0) continue;
            --lv[j];
            bool ok = g.index_of.count(level_key(lv, d)) != 0;
            ++lv[j];
            if (!ok) {
                set_error("build_explicit: level set is not downward closed");
                return kInvalid;
            }
        }
    }
    for (int j = 0; j < d; ++j) g.limits[j] = kMaxLevel;
    fill_offsets(g);
    for (int j = 0; j < d; ++j) g.limits[j] = g.max_level[j];
    if (g.N >= (int64_t(1) << 31)) {
        set_error("build_explicit: grid has 2^31 points or more (offsets are int32)");
        return kRange;
    }
    return kOk;
}

int64_t subspace_id(const Grid& g, const int* levels) {
    int sum = 0;
    for (int j = 0; j < g.d; ++j) {
        if (levels[j] < 0 || levels[j] > g.limits[j]) return -1;
        sum += levels[j];
    }
    if (sum >= g.n) return -1;
    if (g.explicit_set) {
        auto it = g.index_of.find(level_key(levels, g.d));
        return it == g.index_of.end() ? -1 : it->second;
    }
    return g.cnt[sum] + l2idx(g, levels);
}

// rank of a level vector within its level sum: asg.c:266-279
int64_t l2idx(const Grid& g, const int* levels) {
    if (g.explicit_set) {
        int64_t id = subspace_id(g, levels);
        if (id < 0) return -1;
        int sum = 0;
        for (int j = 0; j < g.d; ++j) sum += levels[j];
        return id - g.cnt[sum];
    }
    int64_t rank = 0;
    int sum = levels[0];
    for (int i = 1; i < g.d; ++i) {
        sum += levels[i];
        for (int j = 0; j < levels[i]; ++j) rank += g.dp[size_t(i - 1) * g.n + (sum - j)];
    }
    return rank;
}

// level vector of the given rank within level sum `sum`: asg.c:245-263
int idx2l(const Grid& g, int64_t rank, int sum, int* levels) {
    if (sum < 0 || sum >= g.n || rank < 0) return kRange;
    if (g.explicit_set) {
        int64_t id = g.cnt[sum] + rank;
        if (id >= g.cnt[sum + 1]) return kRange;
        for (int j = 0; j < g.d; ++j) levels[j] = g.sub_levels[id * g.d + j];
        return kOk;
    }
    if (rank >= g.dp[size_t(g.d - 1) * g.n + sum]) return kRange;
    for (int i = g.d - 1; i >= 1; --i) {
        int l = 0;
        while (rank >= g.dp[size_t(i - 1) * g.n + (sum - l)]) {
            rank -= g.dp[size_t(i - 1) * g.n + (sum - l)];
            ++l;
        }
        levels[i] = l;
        sum -= l;
    }
    levels[0] = sum;
    return kOk;
}

// (levels, indices) -> offset: asg.c:187-207
int64_t gp2idx(const Grid& g, const int* levels, const int* indices) {
    int64_t id = subspace_id(g, levels);
    if (id < 0) return -1;
    int64_t in_block = 0;
    for (int j = 0; j < g.d; ++j) {
        if (indices[j] < 0 || indices[j] >= (1 << levels[j])) return -1;
        in_block = (in_block << levels[j]) + indices[j];
    }
    return int64_t(g.sub_base[id]) + in_block;
}

// offset -> (levels, indices): asg.c:210-243
int idx2gp(const Grid& g, int64_t offset, int* levels, int* indices) {
    if (offset < 0 || offset >= g.N) return kRange;
    int s = 0;
    while (s + 1 < g.n && offset >= g.lidx[s + 1]) ++s;
    int64_t r = offset - g.lidx[s];
    int64_t rank = r >> s;
    int64_t bits = r & ((int64_t(1) << s) - 1);
    int st = idx2l(g, rank, s, levels);
    if (st != kOk) return st;
    for (int j = g.d - 1; j >= 0; --j) {
        indices[j] = static_cast<int>(bits & ((int64_t(1) << levels[j]) - 1));
        bits >>= levels[j];
    }
    return kOk;
}

namespace {

// Depth-first walk over the level set with dimension d-1 outermost and dimension 0
// innermost; visit(levels) is called at every leaf.  For a fixed level sum the leaves
// come in rank order, because rank order is lexicographic from dimension d-1 down.
// For an explicit set it also writes the node-code stream the evaluation kernel reads
// (one byte per visited node): at depth J-1 a code below 0x80 marks a full block
// { sum of l_{J-1..0} <= R } and is R; otherwise (code & 0x7f) is the node's bound.
struct DfsWalk {
    const Grid& g;
    int J;
    std::vector<uint8_t>* codes;
    std::vector<int32_t>* leaves;
    // static blocks in walk order: first leaf, budget R, outer level sum
    std::vector<std::array<int32_t, 3>>* blocks = nullptr;
    int64_t nleaf = 0;
    int lv[kMaxDims] = {};
    bool ok = true;

    // largest admissible l_k given the levels fixed above k (and zeros below)
    int bound_at(int k, int rem) {
        if (!g.explicit_set) return std::min(rem, g.limits[k]);
        for (int j = 0; j < k; ++j) lv[j] = 0;
        int b = -1;
        for (int l = 0; l <= std::min(rem, g.limits[k]); ++l) {
            lv[k] = l;
            if (subspace_id(g, lv) < 0) break;
            b = l;
        }
        lv[k] = 0;
        return b;
    }

    // does the subtree under the current prefix hold exactly { sum(l_0..l_k) <= R }?
    int full_block_budget(int k, int rem) {
        int R = -1;
        int64_t count = 0;
        int v[kMaxDims] = {};
        const int J1 = k + 1;
        // enumerate the candidates l_k..l_0 with sum <= rem (odometer, dim 0 fastest)
        while (true) {
            int sum = 0;
            for (int j = 0; j < J1; ++j) { lv[j] = v[j]; sum += v[j]; }
            if (subspace_id(g, lv) >= 0) {
                ++count;
                R = std::max(R, sum);
            }
            int j = 0;
            for (; j < J1; ++j) {
                ++v[j];
                int s2 = 0;
                for (int q = 0; q < J1; ++q) s2 += v[q];
                if (s2 <= rem) break;
                v[j] = 0;
            }
            if (j == J1) break;
        }
        for (int j = 0; j < J1; ++j) lv[j] = 0;
        if (R < 0 || R > 7) return -1;
        int64_t full = 1;
        for (int i = 1; i <= J1; ++i) full = full * (R + i) / i;   // C(R + J, J)
        return count == full ? R : -1;
    }

    void leaf() {
        int64_t id = subspace_id(g, lv);
        if (id < 0) { ok = false; return; }
        if (leaves) leaves->push_back(g.sub_base[id]);
        ++nleaf;
    }

    // budget R of the static block below a depth-J node, or -1: for an explicit set
    // the subtree must be exactly { l_0 + .. + l_{J-1} <= R }; for a restricted-regular
    // set the remaining budget must fit every tabled dimension's limit (the kernel's
    // test in its walk)
    int block_budget(int k, int rem) {
        if (g.explicit_set) return full_block_budget(k, rem);
        for (int j = 0; j < J; ++j)
            if (rem > g.limits[j]) return -1;
        return rem <= 7 ? rem : -1;
    }

    void go(int k, int rem) {
        int b = bound_at(k, rem);
        if (J > 0 && k == J - 1 && (codes || blocks)) {
            int R = block_budget(k, rem);
            if (R >= 0) {
                if (codes) codes->push_back(static_cast<uint8_t>(R));
                if (blocks) {
                    int hsum = 0;
                    for (int j = J; j < g.d; ++j) hsum += lv[j];
                    blocks->push_back({int32_t(nleaf), R, hsum});
                }
                block(k, R);
                return;
            }
            if (codes) codes->push_back(static_cast<uint8_t>(0x80 | b));
        } else if (codes) {
            codes->push_back(static_cast<uint8_t>(b));
        }
        for (int l = 0; l <= b; ++l) {
            lv[k] = l;
            if (k == 0) leaf();
            else go(k - 1, rem - l);
        }
        lv[k] = 0;
    }

    // leaves of a full block in walk order
    void block(int k, int R) {
        for (int l = 0; l <= R; ++l) {
            lv[k] = l;
            if (k == 0) leaf();
            else block(k - 1, R - l);
        }
        lv[k] = 0;
    }
};

}  // namespace

int explicit_tab_dims(const Grid& g) {
    const int J = std::min(g.d, kExplicitTabDims);
    for (int j = 0; j < J; ++j)
        if (g.max_level[j] > 7) return 0;
    return J;
}

namespace {

// Pack the whole static blocks into LDS stages for one element size (see StagePlan).
// LDS bytes of a set of leaves laid out in source order, with the leaves that are
// contiguous in HBM merged into runs and every run starting 16-byte aligned.  leaves
// holds (base offset, elements) pairs, sorted here.
int64_t stage_run_bytes(std::vector<std::pair<int32_t, int64_t>>& leaves, int elem) {
    std::sort(leaves.begin(), leaves.end());
    int64_t bytes = 0, run = 0;
    for (size_t i = 0; i < leaves.size(); ++i) {
        run += leaves[i].second * elem;
        if (i + 1 == leaves.size() || leaves[i].first + leaves[i].second != leaves[i + 1].first) {
            bytes += (run + 15) & ~int64_t(15);
            run = 0;
        }
    }
    return bytes;
}

// Stages hold whole static blocks, in walk order.  Inside a stage the leaves are laid out
// in HBM order: in compact order the subspaces below one node of the walk that share a
// level sum are adjacent, so the leaves of a stage form far fewer contiguous runs than
// leaves, and each run is copied in segments of at most kStageSegBytes.  A stage lists
// its full segments (64 lanes) before its partial ones, so the kernel copies most of it
// without lane masks.
void build_stage_plan(const Grid& g, int J, const std::vector<std::array<int32_t, 3>>& blocks,
                      int elem, StagePlan& p) {
    p = StagePlan();
    if (g.N * elem > int64_t(0x7fffffff)) return;
    // block size of each subspace, by base offset
    std::unordered_map<int32_t, int64_t> size_of;
    size_of.reserve(size_t(g.S) * 2);
    for (int64_t id = 0; id < g.S; ++id) {
        const int64_t end = id + 1 < g.S ? g.sub_base[id + 1] : g.N;
        size_of[g.sub_base[id]] = end - g.sub_base[id];
    }
    p.leaf_lds.assign(size_t(g.S), -1);
    // leaves (walk position) of the open stage, and its buffer
    std::vector<int32_t> open;
    int buf = -1;
    std::vector<std::pair<int32_t, int64_t>> scratch;
    std::vector<int32_t> full, part;                     // segment descriptors of a stage
    auto lay_out = [&]() {
        if (open.empty()) return;
        std::vector<std::pair<int32_t, int32_t>> order;     // (base, walk position)
        order.reserve(open.size());
        for (int32_t q : open) order.emplace_back(g.dfs_base[q], q);
        std::sort(order.begin(), order.end());
        int32_t used = 0;
        size_t i = 0;
        while (i < order.size()) {
            // one run: leaves i..j-1, contiguous in HBM
            size_t j = i + 1;
            while (j < order.size() &&
                   order[j - 1].first + size_of[order[j - 1].first] == order[j].first)
                ++j;
            const int32_t run_base = order[i].first;
            const int32_t run_at = buf * kStageBytes + used;
            for (size_t k = i; k < j; ++k)
                p.leaf_lds[order[k].second] = run_at + (order[k].first - run_base) * elem;
            const int32_t b = int32_t((order[j - 1].first - run_base + size_of[order[j - 1].first]) *
                                      elem);
            for (int32_t o = 0; o < b; o += kStageSegBytes) {
                const int32_t lanes = (std::min(kStageSegBytes, b - o) + 15) / 16;
                auto& to = lanes == kStageSegBytes / 16 ? full : part;
                to.push_back(int32_t(int64_t(run_base) * elem + o));
                to.push_back((run_at + o) | (lanes << kSegLanesShift));
            }
            used += (b + 15) & ~15;
            i = j;
        }
        // the stage's full segments, then its partial ones
        p.seg_begin.push_back(int32_t(p.segs.size() / 2));
        p.segs.insert(p.segs.end(), full.begin(), full.end());
        p.seg_begin.push_back(int32_t(p.segs.size() / 2));
        p.segs.insert(p.segs.end(), part.begin(), part.end());
        full.clear();
        part.clear();
        open.clear();
    };
    for (const auto& blk : blocks) {
        const int32_t q0 = blk[0], R = blk[1], hsum = blk[2];
        if (!stage_whole(J, R, hsum, elem)) continue;
        const int64_t C = stage_binom(R + J, J);
        scratch.clear();
        for (int64_t c = 0; c < C; ++c)
            scratch.emplace_back(g.dfs_base[q0 + c], size_of[g.dfs_base[q0 + c]]);
        if (stage_run_bytes(scratch, elem) > kStageBytes) return;   // the bound is violated
        for (int32_t q : open) scratch.emplace_back(g.dfs_base[q], size_of[g.dfs_base[q]]);
        if (open.empty() || stage_run_bytes(scratch, elem) > kStageBytes) {
            lay_out();                                   // close the stage, open the next
            buf = buf < 0 ? 0 : buf ^ 1;
            p.stage_first.push_back(q0);
        }
        for (int64_t c = 0; c < C; ++c) open.push_back(int32_t(q0 + c));
    }
    lay_out();
    p.stage_first.push_back(int32_t(g.S));
    p.seg_begin.push_back(int32_t(p.segs.size() / 2));
    p.segs.insert(p.segs.end(), 8, 0);                   // 4 empty segments of padding
    p.ok = true;
}

// Static blocks and leaf bases of a regular grid in the sparse order (blk_sparse): the
// host form of the kernel's sparse walk.  blocks: first leaf, budget R, outer level sum.
bool sparse_walk(const Grid& g, int J, std::vector<std::array<int32_t, 3>>& blocks,
                 std::vector<int32_t>& bases) {
    blocks.clear();
    bases.clear();
    int lv[kMaxDims] = {};
    bool ok = true;
    const int nm1 = g.n - 1;
    std::function<void(int, int)> leaves = [&](int k, int R) {     // a static block's leaves
        for (int l = 0; l <= R; ++l) {
            lv[k] = l;
            if (k == 0) {
                const int64_t id = subspace_id(g, lv);
                if (id < 0) ok = false;
                else bases.push_back(g.sub_base[size_t(id)]);
            } else {
                leaves(k - 1, R - l);
            }
        }
        lv[k] = 0;
    };
    std::function<void(int, int)> node = [&](int c, int rem) {
        if (rem > 0)
            for (int k = c - 1; k >= J; --k)
                for (int l = 1; l <= rem; ++l) {
                    lv[k] = l;
                    node(k, rem - l);
                    lv[k] = 0;
                }
        blocks.push_back({int32_t(bases.size()), rem, nm1 - rem});
        if (J > 0) leaves(J - 1, rem);
    };
    node(g.d, nm1);
    return ok && int64_t(bases.size()) == g.S;
}

// The block-staged walk's copy plan (see BlockStagePlan): blocks in walk order, each
// laid out by blk_leaf_offset; a stage takes the next staged blocks while they fit.  The
// kernel applies the same greedy rule with the same compile-time block sizes, so the plan
// needs no per-block table: only the copy segments of each stage.  Every leaf's LDS
// offset is checked here against its HBM layout; a mismatch leaves the plan unusable.
void build_blk_plan(const Grid& g, int J, const std::vector<std::array<int32_t, 3>>& walk_blocks,
                    int elem, int stage_bytes, BlockStagePlan& p) {
    p = BlockStagePlan();
    if (g.explicit_set || g.N * elem > int64_t(0x7fffffff)) return;
    const int nm1 = g.n - 1;
    p.leaf_lds.assign(size_t(g.S), -1);
    for (int j = 0; j < g.d; ++j)
        if (g.limits[j] != nm1) return;                 // the walk assumes a regular grid
    std::vector<std::array<int32_t, 3>> sparse_blocks;
    if (blk_sparse(g.d, elem)) {
        if (!sparse_walk(g, J, sparse_blocks, p.leaf_base)) return;
    } else {
        p.leaf_base = g.dfs_base;
    }
    const std::vector<std::array<int32_t, 3>>& blocks =
        blk_sparse(g.d, elem) ? sparse_blocks : walk_blocks;
    std::vector<int32_t> full, part;
    int used = 0, buf = -1;
    auto close_stage = [&]() {
        if (buf < 0) return;
        p.seg_begin.push_back(int32_t(p.segs.size() / 2));
        p.segs.insert(p.segs.end(), full.begin(), full.end());
        p.seg_begin.push_back(int32_t(p.segs.size() / 2));
        p.segs.insert(p.segs.end(), part.begin(), part.end());
        full.clear();
        part.clear();
        ++p.num_stages;
    };
    auto subspace_of = [&](int32_t base) -> int64_t {
        const auto it = std::lower_bound(g.sub_base.begin(), g.sub_base.end(), base);
        return (it == g.sub_base.end() || *it != base) ? -1 : int64_t(it - g.sub_base.begin());
    };
    // place `bytes` of the next staged unit (a whole block or one slice) in the stages
    auto place = [&](int64_t bytes, int32_t q_first) -> int32_t {
        if (buf < 0 || used + bytes > stage_bytes) {
            close_stage();
            buf = buf < 0 ? 0 : buf ^ 1;
            used = 0;
            p.stage_first.push_back(q_first);
        }
        const int32_t at = buf * stage_bytes + used;
        used += int(bytes);
        return at;
    };
    // lay out leaves [c0, c1) of the block at q0 as a block over the Jb inner dimensions
    // below outer sum Hb (Jb = J: the whole block; Jb = J - 1: one slice) at LDS `at`
    auto lay_out = [&](int32_t q0, int c0, int c1, int Jb, int Hb, int32_t at) -> bool {
        std::vector<std::vector<std::pair<int32_t, int64_t>>> runs(size_t(nm1) + 1);
        int lv[kMaxDims] = {};
        for (int c = c0; c < c1; ++c) {
            const int32_t base = p.leaf_base[size_t(q0 + c)];
            const int64_t id = subspace_of(base);
            if (id < 0) return false;
            int s = 0;
            for (int j = 0; j < Jb; ++j) {
                lv[j] = g.sub_levels[size_t(id * g.d + j)];
                s += lv[j];
            }
            const int64_t lo = blk_leaf_offset(Jb, lv, Hb, elem);
            runs[size_t(s)].emplace_back(base, lo);
            p.leaf_lds[size_t(q0 + c)] = at + int32_t(lo);
        }
        for (int s = 0; s <= nm1; ++s) {
            auto& rl = runs[size_t(s)];
            if (rl.empty()) continue;
            std::sort(rl.begin(), rl.end());
            const int64_t sub_elems = int64_t(1) << (Hb + s);
            const int64_t run_off = blk_run_offset(Jb, s, Hb, elem);
            if (int64_t(rl.size()) != stage_binom(s + Jb - 1, Jb - 1)) return false;
            for (size_t i = 0; i < rl.size(); ++i) {
                if (rl[i].first != rl[0].first + int64_t(i) * sub_elems) return false;   // contiguous
                if (rl[i].second != run_off + int64_t(i) * sub_elems * elem) return false;
            }
            const int32_t len = int32_t(int64_t(rl.size()) * sub_elems * elem);
            const int64_t src = int64_t(rl[0].first) * elem;
            for (int32_t o = 0; o < len; o += kStageSegBytes) {
                const int32_t lanes = (std::min(kStageSegBytes, len - o) + 15) / 16;
                auto& to = lanes == kStageSegBytes / 16 ? full : part;
                to.push_back(int32_t(src + o));
                to.push_back((at + int32_t(run_off) + o) | (lanes << kSegLanesShift));
            }
        }
        return true;
    };
    for (const auto& blk : blocks) {
        const int32_t q0 = blk[0], R = blk[1], H = blk[2];
        if (H != nm1 - R) return;                       // not the regular-walk invariant
        const int how = blk_staging(J, R, H, elem, stage_bytes);
        if (how == 1) {
            const int32_t at = place(blk_bytes(J, R, H, elem), q0);
            if (!lay_out(q0, 0, int(stage_binom(R + J, J)), J, H, at)) return;
        } else if (how == 2) {
            // slices a = 0..R: leaves with l_{J-1} = a, consecutive in walk order
            int c = 0;
            for (int a = 0; a <= R; ++a) {
                const int cnt = int(stage_binom(R - a + J - 1, J - 1));
                const int32_t at = place(blk_bytes(J - 1, R - a, H + a, elem), q0 + c);
                if (!lay_out(q0, c, c + cnt, J - 1, H + a, at)) return;
                c += cnt;
            }
        }
    }
    close_stage();
    p.stage_first.push_back(int32_t(g.S));
    p.seg_begin.push_back(int32_t(p.segs.size() / 2));
    p.segs.insert(p.segs.end(), 8, 0);                   // 4 empty segments of padding
    p.ok = true;
}

}  // namespace

int build_launch_tables(Grid& g, int hier_tile_values) {
    const int d = g.d;
    // evaluation walk
    g.dfs_base.clear();
    g.dfs_base.reserve(g.S);
    g.dfs_bound.clear();
    DfsWalk w{g, g.explicit_set ? explicit_tab_dims(g) : 0,
              g.explicit_set ? &g.dfs_bound : nullptr, &g.dfs_base};
    w.go(g.d - 1, g.n - 1);
    if (!w.ok || int64_t(g.dfs_base.size()) != g.S) {
        set_error("build_launch_tables: depth-first walk does not cover the level set");
        return kInvalid;
    }
    // staged walk: the static blocks of the walk with kStageJ tabled dimensions (levels
    // of the tabled dimensions below 8), packed into LDS stages
    g.stage[0] = StagePlan();
    g.stage[1] = StagePlan();
    g.blk[0] = BlockStagePlan();
    g.blk[1] = BlockStagePlan();
    {
        const int J = g.explicit_set ? explicit_tab_dims(g) : std::min(d, kStageJ);
        bool tabled = J > 0;
        for (int j = 0; j < J; ++j) tabled = tabled && g.max_level[j] <= 7 && g.limits[j] <= 7;
        if (tabled) {
            std::vector<std::array<int32_t, 3>> blocks;
            DfsWalk bw{g, J, nullptr, nullptr};
            bw.blocks = &blocks;
            bw.go(d - 1, g.n - 1);
            if (bw.ok && bw.nleaf == g.S) {
                build_stage_plan(g, J, blocks, 8, g.stage[0]);
                build_stage_plan(g, J, blocks, 4, g.stage[1]);
                if (J == kStageJ) {
                    build_blk_plan(g, J, blocks, 8, blk_stage_bytes(d, g.n - 1, 8), g.blk[0]);
                    build_blk_plan(g, J, blocks, 4, blk_stage_bytes(d, g.n - 1, 4), g.blk[1]);
                }
            }
        }
    }
    // hierarchization: pole families per dimension
    g.hier.assign(d, HierTables());
    int lv[kMaxDims];
    for (int k = 0; k < d; ++k) {
        HierTables& h = g.hier[k];
        int32_t tile_first = 0, tile_fams = 0, tile_vals = 0;
        auto close_tile = [&]() {
            if (tile_fams == 0) return;
            h.tiles.push_back(tile_first);
            h.tiles.push_back(tile_fams);
            h.tiles.push_back(tile_vals);
            h.max_families_per_tile = std::max(h.max_families_per_tile, int(tile_fams));
            tile_first += tile_fams;
            tile_fams = 0;
            tile_vals = 0;
        };
        for (int64_t id = 0; id < g.S; ++id) {
            const uint8_t* l = &g.sub_levels[id * d];
            if (l[k] != 0) continue;               // families are rooted at l_k = 0
            int c = 0, B = 0;
            for (int j = 0; j < d; ++j) {
                lv[j] = l[j];
                c += l[j];
                if (j > k) B += l[j];
            }
            std::vector<int32_t> bases;
            for (int m = 0;; ++m) {
                lv[k] = m;
                int64_t sid = subspace_id(g, lv);
                if (sid < 0) break;
                bases.push_back(g.sub_base[sid]);
            }
            int m = int(bases.size()) - 1;
            if (m == 0) continue;                   // l_k = 0 points have no parents in k
            int64_t vals = (int64_t(1) << c) * ((int64_t(2) << m) - 1);
            if (vals > (int64_t(1) << 24)) {
                set_error("build_launch_tables: pole family too large");
                return kUnsupported;
            }
            h.max_family_values = std::max<int>(h.max_family_values, int(vals));
            if (tile_fams > 0 && tile_vals + vals > hier_tile_values) close_tile();
            h.fams.push_back(tile_vals);
            h.fams.push_back(m | (c << 8) | (B << 16));
            h.fams.push_back(int32_t(h.blocks.size()));
            h.blocks.insert(h.blocks.end(), bases.begin(), bases.end());
            tile_vals += int32_t(vals);
            ++tile_fams;
        }
        close_tile();
    }
    // hierarchization: plane families of the dimension pairs (2p, 2p+1)
    g.hier2.assign(d / 2, PairTables());
    for (int p = 0; p < d / 2; ++p) {
        const int k = 2 * p;
        PairTables& h = g.hier2[p];
        const int tile_target = std::min(std::max(hier_tile_values, 64), kPairMaxValues);
        int32_t tile_first = 0, tile_fams = 0, tile_vals = 0, tile_subs = 0, tile_blk0 = 0;
        auto close_tile = [&]() {
            if (tile_fams == 0) return;
            h.tiles.push_back(tile_first);
            h.tiles.push_back(tile_fams);
            h.tiles.push_back(tile_vals);
            h.tiles.push_back(h.fams[3 * tile_first + 2]);
            h.tiles.push_back(tile_subs);
            h.tiles.push_back(tile_blk0);
            h.tiles.push_back(int32_t(h.blocks.size()) - tile_blk0);
            h.tiles.push_back(0);
            h.max_families_per_tile = std::max(h.max_families_per_tile, int(tile_fams));
            h.max_subs_per_tile = std::max(h.max_subs_per_tile, int(tile_subs));
            h.max_blocks_per_tile =
                std::max(h.max_blocks_per_tile, int(h.blocks.size()) - int(tile_blk0));
            h.max_tile_values = std::max(h.max_tile_values, int(tile_vals));
            tile_first += tile_fams;
            tile_fams = 0;
            tile_vals = 0;
            tile_subs = 0;
        };
        bool fits = true;
        for (int64_t id = 0; id < g.S && fits; ++id) {
            const uint8_t* l = &g.sub_levels[id * d];
            if (l[k] != 0 || l[k + 1] != 0) continue;     // families rooted at (0, 0)
            int c = 0, Bfix = 0;
            for (int j = 0; j < d; ++j) {
                lv[j] = l[j];
                if (j != k && j != k + 1) c += l[j];
                if (j > k + 1) Bfix += l[j];
            }
            std::vector<int32_t> sub_first, sub_block, bases;
            int64_t vals = 0;
            int nblocks = 0;
            for (int a = 0;; ++a) {
                lv[k] = a;
                lv[k + 1] = 0;
                if (subspace_id(g, lv) < 0) break;
                sub_first.push_back(int32_t(vals));
                sub_block.push_back(int32_t(bases.size()));
                int mb = -1;
                for (int b = 0;; ++b) {
                    lv[k + 1] = b;
                    const int64_t sid = subspace_id(g, lv);
                    if (sid < 0) break;
                    bases.push_back(g.sub_base[sid]);
                    mb = b;
                }
                vals += (int64_t(1) << (c + a)) * ((int64_t(2) << mb) - 1);
                nblocks += mb + 1;
                if (vals > kPairMaxValues) break;
            }
            lv[k] = 0;
            lv[k + 1] = 0;
            if (vals > kPairMaxValues) {
                fits = false;
                break;
            }
            if (nblocks == 1) continue;                    // one subspace: no parents
            if (tile_fams > 0 && tile_vals + vals > tile_target) close_tile();
            if (tile_fams == 0) tile_blk0 = int32_t(h.blocks.size());
            h.fams.push_back(tile_vals);
            h.fams.push_back(c | (Bfix << 8) | ((int(sub_first.size()) - 1) << 16));
            h.fams.push_back(int32_t(h.subs.size() / 2));
            for (size_t a = 0; a < sub_first.size(); ++a) {
                h.subs.push_back(sub_first[a]);
                h.subs.push_back(int32_t(h.blocks.size()) + sub_block[a]);
            }
            h.blocks.insert(h.blocks.end(), bases.begin(), bases.end());
            tile_vals += int32_t(vals);
            tile_subs += int32_t(sub_first.size());
            ++tile_fams;
        }
        close_tile();
        h.ok = fits;
        if (!fits) h = PairTables();
    }
    return kOk;
}

}  // namespace sg
