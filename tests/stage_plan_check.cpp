// Host check of the staged evaluation's copy plan (StagePlan in sg_grid.hpp).  It runs
// the plan's segment copies on a byte image of the surplus array and checks, per stage:
//   * seg_begin lists the stage's full segments (64 lanes) and then its partial ones;
//   * segments are 16-byte aligned in LDS, stay inside the stage's buffer (k % 2) and do
//     not overlap, so the waves' copies never race on a byte;
//   * every staged leaf of the stage finds its own surplus bytes at leaf_lds;
//   * stages start at increasing leaves and every staged leaf lies in one stage.
// The block-staged walk's plan (BlockStagePlan) is checked the same way against the
// kernel's own addressing, re-derived here: the walk's static blocks in depth-first
// order, each staged while blk_bytes fits the stage, stages filled greedily, every leaf
// at its block's base plus blk_leaf_offset.
// Built and run by tests/test_stage_plan.py; prints one line per grid and element size.
#include "sg_grid.hpp"

#include <algorithm>
#include <cstdio>
#include <vector>

namespace {

int fail(const char* what, int d, int elem, int k) {
    std::printf("FAIL d=%d elem=%d stage %d: %s\n", d, elem, k, what);
    return 1;
}

int check_plan(const sg::Grid& g, int e) {
    const int elem = e ? 4 : 8;
    const sg::StagePlan& p = g.stage[e];
    if (!p.ok) {
        std::printf("d=%d n=%d elem=%d: no plan\n", g.d, g.n, elem);
        return 0;
    }
    const int K = int(p.stage_first.size()) - 1;
    if (int(p.seg_begin.size()) != 2 * K + 1) return fail("seg_begin has 2K+1 entries", g.d, elem, -1);
    if (int(p.segs.size()) != 2 * (p.seg_begin[2 * K] + 4))
        return fail("descriptor table ends in four empty entries", g.d, elem, -1);
    std::vector<uint8_t> src(size_t(g.N) * elem);
    for (size_t i = 0; i < src.size(); ++i) src[i] = uint8_t((i * 2654435761u) >> 13);
    std::vector<int64_t> size_at(size_t(g.N), 0);      // elements of the subspace at a base
    for (int64_t id = 0; id < g.S; ++id)
        size_at[size_t(g.sub_base[id])] = (id + 1 < g.S ? g.sub_base[id + 1] : g.N) - g.sub_base[id];
    int64_t full = 0, partial = 0, staged = 0;
    for (int k = 0; k < K; ++k) {
        if (p.stage_first[k] >= p.stage_first[k + 1]) return fail("stage order", g.d, elem, k);
        if (p.seg_begin[2 * k] > p.seg_begin[2 * k + 1] || p.seg_begin[2 * k + 1] > p.seg_begin[2 * k + 2])
            return fail("seg_begin order", g.d, elem, k);
        std::vector<uint8_t> lds(2 * sg::kStageBytes, 0);
        std::vector<char> written(2 * sg::kStageBytes, 0);
        for (int s = p.seg_begin[2 * k]; s < p.seg_begin[2 * k + 2]; ++s) {
            const int64_t from = p.segs[2 * s];
            const int at = p.segs[2 * s + 1] & sg::kSegLdsMask,
                      lanes = p.segs[2 * s + 1] >> sg::kSegLanesShift;
            const bool is_full = s < p.seg_begin[2 * k + 1];
            if (is_full != (lanes == 64)) return fail("full segments first", g.d, elem, k);
            if (lanes < 1 || at % 16) return fail("segment shape", g.d, elem, k);
            if (at / sg::kStageBytes != k % 2 || (at + 16 * lanes - 1) / sg::kStageBytes != k % 2)
                return fail("segment outside the stage's buffer", g.d, elem, k);
            for (int b = 0; b < 16 * lanes; ++b) {
                if (written[at + b]) return fail("segments overlap", g.d, elem, k);
                written[at + b] = 1;
                const int64_t o = from + b;        // bytes past the array read as zero
                lds[at + b] = o < int64_t(src.size()) ? src[size_t(o)] : 0;
            }
            (is_full ? full : partial) += 1;
        }
        for (int q = p.stage_first[k]; q < p.stage_first[k + 1]; ++q) {
            const int32_t at = p.leaf_lds[q];
            if (at < 0) continue;
            ++staged;
            const int64_t base = g.dfs_base[q], bytes = size_at[size_t(base)] * elem;
            if (at % elem) return fail("leaf alignment", g.d, elem, k);
            for (int64_t b = 0; b < bytes; ++b)
                if (!written[at + b] || lds[at + b] != src[size_t(base * elem + b)])
                    return fail("leaf bytes differ from its surpluses", g.d, elem, k);
        }
    }
    for (int q = 0; q < (K ? p.stage_first[0] : int(g.S)); ++q)
        if (p.leaf_lds[q] >= 0) return fail("staged leaf before the first stage", g.d, elem, -1);
    std::printf("d=%d n=%d elem=%d: %d stages, %lld full + %lld partial segments, %lld of %lld leaves staged: ok\n",
                g.d, g.n, elem, K, (long long)full, (long long)partial, (long long)staged,
                (long long)g.S);
    return 0;
}

// the static blocks of a regular grid's walk, in the kernel's visiting order: outer
// levels l_{d-1..J}
struct BlkWalk {
    const sg::Grid& g;
    int J;
    std::vector<std::vector<int>> blocks;   // outer levels (index k for dims J..d-1)
    std::vector<int> lv;
    // depth first: dimension d-1 outermost, levels ascending
    void go(int k, int rem) {
        if (k == J - 1) {
            blocks.push_back(lv);
            return;
        }
        for (int l = 0; l <= rem; ++l) {
            lv[k] = l;
            go(k - 1, rem - l);
        }
        lv[k] = 0;
    }
    // sparse: below a node with dimensions >= c fixed, the children l_k = 1..rem for
    // k = c-1 down to J, then the node's own block (dimensions J..c-1 at level 0)
    void sparse(int c, int rem) {
        for (int k = c - 1; k >= J && rem > 0; --k)
            for (int l = 1; l <= rem; ++l) {
                lv[k] = l;
                sparse(k, rem - l);
                lv[k] = 0;
            }
        blocks.push_back(lv);
    }
};

int check_blk_plan(const sg::Grid& g, int e) {
    const int elem = e ? 4 : 8;
    const sg::BlockStagePlan& p = g.blk[e];
    if (!p.ok) {
        std::printf("d=%d n=%d elem=%d: no block plan\n", g.d, g.n, elem);
        return 0;
    }
    const int J = std::min(g.d, sg::kStageJ), nm1 = g.n - 1, STB = sg::blk_stage_bytes(g.d, g.n - 1, elem);
    const int K = p.num_stages;
    if (int(p.seg_begin.size()) != 2 * K + 1) return fail("blk seg_begin has 2K+1 entries", g.d, elem, -1);
    if (int(p.segs.size()) != 2 * (p.seg_begin[2 * K] + 4))
        return fail("blk descriptor table ends in four empty entries", g.d, elem, -1);
    std::vector<uint8_t> src(size_t(g.N) * elem);
    for (size_t i = 0; i < src.size(); ++i) src[i] = uint8_t((i * 2654435761u) >> 11);
    BlkWalk w{g, J, {}, std::vector<int>(g.d, 0)};
    if (sg::blk_sparse(g.d, elem)) w.sparse(g.d, nm1);
    else w.go(g.d - 1, nm1);
    // the kernel's stage assignment
    struct Leaf { int64_t base, bytes; int lds; };
    std::vector<std::vector<Leaf>> stage_leaves;
    int used = 0, buf = -1;
    int64_t staged = 0, total = 0, sliced = 0;
    auto place = [&](int64_t bytes) -> int {
        if (buf < 0 || used + bytes > STB) {
            buf = buf < 0 ? 0 : buf ^ 1;
            used = 0;
            stage_leaves.emplace_back();
        }
        const int at = buf * STB + used;
        used += int(bytes);
        return at;
    };
    for (const auto& outer : w.blocks) {
        int H = 0;
        for (int k = J; k < g.d; ++k) H += outer[k];
        const int R = nm1 - H;
        const int64_t C = sg::stage_binom(R + J, J);
        total += C;
        const int how = sg::blk_staging(J, R, H, elem, STB);
        if (how == 0) continue;
        // leaves of the block (inner levels with sum <= R) in walk order, dim J-1 slowest
        std::vector<std::vector<int>> leaves;
        std::vector<int> v(J, 0);
        for (;;) {
            int s = 0;
            for (int j = 0; j < J; ++j) s += v[j];
            if (s <= R) leaves.push_back(v);
            int j = 0;
            while (j < J && ++v[j] > R) v[j++] = 0;
            if (j == J) break;
        }
        auto add = [&](const std::vector<int>& inner_v, int Jb, int Hb, int at) -> int {
            std::vector<int> l(outer);
            int s = 0, inner[sg::kMaxDims] = {};
            for (int j = 0; j < J; ++j) l[j] = inner_v[j];
            for (int j = 0; j < Jb; ++j) { inner[j] = inner_v[j]; s += inner_v[j]; }
            const int64_t id = sg::subspace_id(g, l.data());
            if (id < 0) return 1;
            const int lds = at + int(sg::blk_leaf_offset(Jb, inner, Hb, elem));
            stage_leaves.back().push_back({g.sub_base[id], (int64_t(1) << (Hb + s)) * elem, lds});
            ++staged;
            return 0;
        };
        if (how == 1) {
            const int at = place(sg::blk_bytes(J, R, H, elem));
            for (const auto& lv : leaves)
                if (add(lv, J, H, at)) return fail("blk leaf not in the grid", g.d, elem, -1);
        } else {
            ++sliced;
            for (int a = 0; a <= R; ++a) {
                const int at = place(sg::blk_bytes(J - 1, R - a, H + a, elem));
                for (const auto& lv : leaves)
                    if (lv[J - 1] == a && add(lv, J - 1, H + a, at))
                        return fail("blk leaf not in the grid", g.d, elem, -1);
            }
        }
    }
    if (total != g.S) return fail("blk walk does not cover the grid", g.d, elem, -1);
    if (int(stage_leaves.size()) != K) return fail("blk stage count", g.d, elem, -1);
    int64_t full = 0, partial = 0;
    for (int k = 0; k < K; ++k) {
        std::vector<uint8_t> lds(2 * STB, 0);
        std::vector<char> written(2 * STB, 0);
        for (int s = p.seg_begin[2 * k]; s < p.seg_begin[2 * k + 2]; ++s) {
            const int64_t from = p.segs[2 * s];
            const int at = p.segs[2 * s + 1] & sg::kSegLdsMask,
                      lanes = p.segs[2 * s + 1] >> sg::kSegLanesShift;
            const bool is_full = s < p.seg_begin[2 * k + 1];
            if (is_full != (lanes == 64)) return fail("blk full segments first", g.d, elem, k);
            if (lanes < 1 || at % 16) return fail("blk segment shape", g.d, elem, k);
            if (at / STB != k % 2 || (at + 16 * lanes - 1) / STB != k % 2)
                return fail("blk segment outside the stage's buffer", g.d, elem, k);
            for (int b = 0; b < 16 * lanes; ++b) {
                if (written[at + b]) return fail("blk segments overlap", g.d, elem, k);
                written[at + b] = 1;
                const int64_t o = from + b;
                lds[at + b] = o < int64_t(src.size()) ? src[size_t(o)] : 0;
            }
            (is_full ? full : partial) += 1;
        }
        for (const Leaf& lf : stage_leaves[k]) {
            if (lf.lds % elem) return fail("blk leaf alignment", g.d, elem, k);
            for (int64_t b = 0; b < lf.bytes; ++b)
                if (!written[lf.lds + b] || lds[lf.lds + b] != src[size_t(lf.base * elem + b)])
                    return fail("blk leaf bytes differ from its surpluses", g.d, elem, k);
        }
    }
    std::printf("d=%d n=%d elem=%d block-staged (%s order): %d stages, %lld full + %lld partial segments, %lld of %lld leaves staged (%lld blocks by slices): ok\n",
                g.d, g.n, elem, sg::blk_sparse(g.d, elem) ? "sparse" : "depth-first", K,
                (long long)full, (long long)partial, (long long)staged, (long long)g.S,
                (long long)sliced);
    return 0;
}

int check_regular(int d, int n, const int* limits) {
    sg::Grid g;
    if (sg::build_regular(g, d, n, limits) != 0 || sg::build_launch_tables(g, 512) != 0) {
        std::printf("FAIL d=%d n=%d: grid\n", d, n);
        return 1;
    }
    return check_plan(g, 0) | check_plan(g, 1) | check_blk_plan(g, 0) | check_blk_plan(g, 1);
}

// the regular level set of dimension d and level sums below n, plus l_0 up to top
int check_explicit(int d, int n, int top) {
    std::vector<uint8_t> lv;
    std::vector<int> l(d, 0);
    for (;;) {
        int sum = 0;
        for (int v : l) sum += v;
        bool axis = true;
        for (int j = 1; j < d; ++j) axis = axis && l[j] == 0;
        if (sum < n || (axis && l[0] <= top))
            for (int v : l) lv.push_back(uint8_t(v));
        int j = 0;
        while (j < d && ++l[j] > (j == 0 ? std::max(n, top) : n)) l[j++] = 0;
        if (j == d) break;
    }
    sg::Grid g;
    if (sg::build_explicit(g, d, lv.data(), int64_t(lv.size() / d)) != 0 ||
        sg::build_launch_tables(g, 512) != 0) {
        std::printf("FAIL explicit d=%d: grid\n", d);
        return 1;
    }
    return check_plan(g, 0) | check_plan(g, 1);
}

}  // namespace

int main() {
    int r = 0;
    r |= check_regular(10, 8, nullptr);
    r |= check_regular(12, 8, nullptr);
    r |= check_regular(20, 6, nullptr);
    r |= check_regular(1, 8, nullptr);
    r |= check_regular(2, 8, nullptr);
    r |= check_regular(3, 5, nullptr);
    r |= check_regular(5, 9, nullptr);
    r |= check_regular(16, 4, nullptr);
    const int lim6[6] = {3, 1, 3, 2, 6, 6};
    r |= check_regular(6, 7, lim6);
    const int lim3[3] = {1, 5, 5};
    r |= check_regular(3, 6, lim3);
    r |= check_explicit(4, 4, 6);
    r |= check_explicit(6, 5, 7);
    std::printf(r ? "FAILED\n" : "ALL OK\n");
    return r;
}
