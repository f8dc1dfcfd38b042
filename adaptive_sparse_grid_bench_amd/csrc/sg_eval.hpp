// Evaluation kernel: the interpolant u(x) = sum over subspaces l of prod_k hat_k(l_k) *
// surplus[base(l) + in-block index], one query point per lane (values of
// adaptive_sparse_grid_bench.c "asg.c":1237-1274).
//
// Instead of recomputing the d-term hat product for every subspace, each lane walks the
// level set depth first (dimension d-1 outermost, dimension 0 innermost) and keeps
// prefix products, so a leaf costs one multiply, one shift-add and one FMA.  Leaf
// base offsets come from a table in walk order, read through the scalar cache (the
// walk is uniform across the wave).  The J innermost dimensions keep hats and offsets
// for every level in register tables; the leaves below a depth-J node then form a
// static block whose gathers are issued in batches before any is consumed.
//
// Template parameters:
//   D    dimensions;  LT  levels per register table (8) or 16 (no tables);
//   J    tabled dimensions;  B  gathers per batch;  W  waves per SIMD the register
//        allocation targets (0 = compiler's choice);
//   T    coordinate / surplus type;  Acc  accumulator (FP64 FMA or FP32 Kahan);
//   EXPL explicit (refined-subspace) level set: node bounds come from a byte stream;
//   STG  staged: the workgroup copies the surplus blocks of the coming static blocks
//        into LDS (buffer loads straight to LDS, 16 bytes per lane, double-buffered
//        stages planned on the host) and each lane reads its coefficient there.  Every
//        lane of a workgroup visits the same leaves in the same order, so one copy
//        serves all its waves; a direct gather costs one texture-address pass per wave
//        and leaf instead.  Blocks too large to share a stage, and leaves outside static
//        blocks, are still gathered from global memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "sg_grid.hpp"
#include "sg_kernels.hpp"

namespace sg {
namespace eval {

constexpr int kThreads = 256;
static_assert(kThreads == kBlkThreads, "work items fill one block-staged workgroup");

template <class T>
struct Params {
    const T* alpha;
    const T* coords;            // row-major [P][D] per grid
    T* out;
    const GridDesc* grids;
    const WorkItem* work;       // null: one grid (grids[0]), workgroups tile [0, P)
    const uint32_t* perm;       // one-grid launches: lane order of the points, or null
    int64_t P;                  // one-grid launches: number of points
};

// pin(): the sum so far must exist at this point of the program (an empty asm that
// reads and writes it), so the compiler cannot sink a batch's accumulation below a
// later branch and keep that batch's loaded values alive across it.
struct AccF64 {
    double s = 0.0;
    __device__ __forceinline__ void add(double w, double a) { s = fma(w, a, s); }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ void pin() { asm volatile("" : "+v"(s)); }
};

// FP64 sums a static block nested (block_consume_nested); the Kahan accumulator keeps
// its leaf-by-leaf compensated sum.
template <class Acc>
inline constexpr bool kNestedSum = false;
template <>
inline constexpr bool kNestedSum<AccF64> = true;

// Kahan-compensated FP32 accumulation across subspaces.  Contraction is off inside the
// update, so every walk variant computes exactly the written compensation whatever the
// compiler knows about s and c (at the first leaf both are constant zeros).
struct AccKahan {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float w, float a) {
#pragma clang fp contract(off)
        const float y = fmaf(w, a, -c);
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s; }
    __device__ __forceinline__ void pin() { asm volatile("" : "+v"(s), "+v"(c)); }
};

// Hat values and in-block byte offsets of one coordinate at levels 0..LT-1.
// With m_l = x 2^(l+1) - (2 t_l + 1), t_l = floor(x 2^l):  hat_l = 1 - |m_l|,
// t_{l+1} = 2 t_l + [m_l >= 0],  m_{l+1} = 2 m_l - (m_l >= 0 ? 1 : -1).  For x on the
// 2^-24 lattice every step is exact, so hat_l equals asg.c:1257-1262's value.
template <int LT, class T>
__device__ __forceinline__ void level_table(T x, T (&ph)[LT], uint32_t (&ub)[LT]) {
    T m = T(2) * x - T(1);
    uint32_t t = 0;
#pragma unroll
    for (int l = 0; l < LT; ++l) {
        ph[l] = T(1) - fabs(m);
        ub[l] = t * uint32_t(sizeof(T));
        const bool ge = m >= T(0);
        t = 2u * t + (ge ? 1u : 0u);
        m = T(2) * m + (ge ? T(-1) : T(1));
    }
}

// Surplus gathers are raw buffer loads: the subspace's base offset rides in the
// instruction's scalar offset and the lane's in-block byte offset in its vector offset,
// so a leaf needs no 64-bit address arithmetic; reads past N return 0.
__device__ __forceinline__ double gather(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                         double*) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float gather(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                        float*) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

template <class T>
__device__ __forceinline__ T gather_leaf(__amdgpu_buffer_rsrc_t r, int32_t base, uint32_t voff) {
    return gather(r, voff, uint32_t(base) * uint32_t(sizeof(T)), static_cast<T*>(nullptr));
}

// Read-only tables (work items, grid descriptors, walk-order leaf bases, node codes)
// are read through the constant address space: with a wave-uniform address the loads
// become scalar loads, so a gather's subspace base stays in a scalar register.
template <class V>
using cptr = const __attribute__((address_space(4))) V*;

template <class V>
__device__ __forceinline__ cptr<V> as_const(const V* p) { return (cptr<V>)p; }

using lds_byte = __attribute__((address_space(3))) uint8_t;

// s_waitcnt immediate of the gfx9 family for vmcnt(0) alone: vmcnt bits 3:0 and 15:14
// zero, expcnt (6:4) and lgkmcnt (11:8) at their maxima
constexpr int kWaitVmcnt0 = 0x0F70;

// Copy stage k into LDS, 16 bytes per lane and segment: each wave takes a contiguous
// share of the stage's full segments (64 lanes, no lane mask) and of its partial ones,
// and reads their descriptors four at a time (the descriptor table ends in four empty
// entries).  Segment g is (source byte offset, LDS byte offset | lanes << kSegLanesShift).
template <int TPB>
__device__ __forceinline__ void copy_stage(__amdgpu_buffer_rsrc_t alpha, lds_byte* lds,
                                           cptr<int32_t> seg_begin, cptr<int32_t> segs, int k) {
    const int s0 = seg_begin[2 * k], sp = seg_begin[2 * k + 1], s1 = seg_begin[2 * k + 2];
    constexpr int kWaves = TPB / 64;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    int per = (sp - s0 + kWaves - 1) / kWaves;
    int g0 = s0 + wave * per;
    int g1 = min(sp, g0 + per);
    int g = g0;
    for (; g + 4 <= g1; g += 4) {
        int v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = segs[2 * g + i];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(
                alpha, lds + uint32_t(v[2 * j + 1] & kSegLdsMask), 16, lane * 16u, v[2 * j], 0, 0);
    }
    for (; g < g1; ++g)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            alpha, lds + uint32_t(segs[2 * g + 1] & kSegLdsMask), 16, lane * 16u, segs[2 * g], 0, 0);
    per = (s1 - sp + kWaves - 1) / kWaves;
    g0 = sp + wave * per;
    g1 = min(s1, g0 + per);
    for (g = g0; g < g1; g += 4) {
        int v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = segs[2 * g + i];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (g + j < g1 && int(lane) < (v[2 * j + 1] >> kSegLanesShift))
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    alpha, lds + uint32_t(v[2 * j + 1] & kSegLdsMask), 16, lane * 16u, v[2 * j], 0, 0);
        }
    }
}

// a stage barrier: this wave's copies into LDS have landed (buffer loads to LDS count in
// vmcnt, and the workgroup barrier's fence does not wait for them), then every wave of
// the workgroup has arrived
__device__ __forceinline__ void stage_barrier() {
    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();
}

template <class T, class Acc, bool STG>
struct State {
    static constexpr bool kStaged = STG;
    static constexpr bool kBlk = false;
    __amdgpu_buffer_rsrc_t alpha;
    cptr<int32_t> leaf_base;
    cptr<uint8_t> bounds;
    int q;                      // next leaf (uniform)
    int bq;                     // next node code (uniform; explicit sets)
    Acc acc;
    // staged walk (uniform across the workgroup)
    lds_byte* lds;              // two stage buffers of kStageBytes
    cptr<int32_t> leaf_lds, stage_first, seg_begin, segs;
    int stage, num_stages;
    int next_first;             // first leaf of the next stage

    __device__ __forceinline__ void issue(int k) {
        copy_stage<kThreads>(alpha, lds, seg_begin, segs, k);
    }
    // every wave is done with the current stage and the next one has landed: move on
    // and start copying the one after into the buffer just released
    __device__ __forceinline__ void advance() {
        stage_barrier();
        ++stage;
        next_first = stage_first[stage + 1];
        if (stage + 1 < num_stages) issue(stage + 1);
    }
    // a staged static block starts at leaf q0
    __device__ __forceinline__ void block_start(int q0) {
        acc.pin();
        if (q0 == next_first) advance();
    }
    __device__ __forceinline__ T fetch_lds(int qq, uint32_t voff) const {
        return *reinterpret_cast<const __attribute__((address_space(3))) T*>(
            lds + (uint32_t(leaf_lds[qq]) + voff));
    }
    __device__ __forceinline__ T fetch_global(int qq, uint32_t voff) const {
        return gather_leaf<T>(alpha, leaf_base[qq], voff);
    }
    __device__ __forceinline__ void leaf(T w, uint32_t voff) {
        acc.add(w, fetch_global(q, voff));
        ++q;
    }
};

// State of the block-staged walk (grids whose walk invariant H = n - 1 - R holds, n - 1
// = NM1 a compile-time constant).  Staged blocks follow each other in the current stage
// buffer; cur is where the next one goes.  A block that does not fit what is left of the
// buffer starts the next stage (the host planned the stages by the same rule).
template <class T, class Acc, int NM1, int TPB, int STB>
struct StateBlk {
    static constexpr bool kStaged = true;
    static constexpr bool kBlk = true;
    static constexpr int kNM1 = NM1;
    static constexpr int kStageBytes = STB;
    __amdgpu_buffer_rsrc_t alpha;
    cptr<int32_t> leaf_base;
    cptr<uint8_t> bounds;
    int q;                      // next leaf (uniform)
    int bq;                     // unused (explicit sets do not take this walk)
    Acc acc;
    lds_byte* lds;              // two stage buffers of STB bytes
    cptr<int32_t> seg_begin, segs;
    uint32_t base, cur, limit;  // LDS byte offsets: current block, next block, buffer end
    int stage, num_stages;

    __device__ __forceinline__ void issue(int k) {
        copy_stage<TPB>(alpha, lds, seg_begin, segs, k);
    }
    // stage 0 landed and stage 1 on its way; called once, before the walk
    __device__ __forceinline__ void begin() {
        if (num_stages > 0) {
            issue(0);
            stage_barrier();
            if (num_stages > 1) issue(1);
        }
    }
    __device__ __forceinline__ void advance() {
        stage_barrier();
        ++stage;
        cur = uint32_t(stage & 1) * uint32_t(STB);
        limit = cur + uint32_t(STB);
        if (stage + 1 < num_stages) issue(stage + 1);
    }
    template <uint32_t BYTES>
    __device__ __forceinline__ void block_begin() {
        acc.pin();
        if (cur + BYTES > limit) advance();
        base = cur;
        cur += BYTES;
    }
    __device__ __forceinline__ T fetch_lds_at(uint32_t a) const {
        return *reinterpret_cast<const __attribute__((address_space(3))) T*>(lds + a);
    }
    __device__ __forceinline__ T fetch_global(int qq, uint32_t voff) const {
        return gather_leaf<T>(alpha, leaf_base[qq], voff);
    }
    __device__ __forceinline__ void leaf(T w, uint32_t voff) {
        acc.add(w, fetch_global(q, voff));
        ++q;
    }
};

// Per-lane state: m0 = 2x - 1 per dimension for the incremental walk, and the hat /
// offset tables of the J innermost dimensions.  PACK (the FP32 block-staged walk at 4
// waves per SIMD, blk_quad): the packed tables of LaneBlk (ub_at, ph_at); 128 VGPRs and
// 44 bytes of scratch per lane at d = 10, n = 8.
template <int D, int LT, int J, class T, bool PACK = false>
struct Lane {
    static constexpr bool kPackedUb = PACK;
    static constexpr bool kTopOnTheFly = PACK && J > 1;
    static constexpr int kJ = J;
    static constexpr int kElem = int(sizeof(T));
    T m0[D];
    T ph[J > 0 ? J : 1][J > 0 ? LT : 1];
    uint32_t ub[J > 0 && !PACK ? J : 1][J > 0 ? LT : 1];
    uint32_t t7[J > 0 ? J : 1];     // PACK: t_7 of each tabled dimension
    T x[J > 0 ? J : 1];             // PACK: the outermost tabled dimension's coordinate
};

// The level vectors (l_{J-1}, ..., l_0) with sum <= R in walk order (dim J-1
// outermost), built at compile time.
__host__ __device__ constexpr int binom(int a, int b) {
    int r = 1;
    for (int i = 1; i <= b; ++i) r = r * (a - b + i) / i;
    return r;
}

template <int J, int R>
struct Simplex {
    static constexpr int C = binom(R + J, J);
    int l[C][J];
    __host__ __device__ constexpr Simplex() : l{} {
        int v[J] = {};
        for (int c = 0; c < C; ++c) {
            for (int j = 0; j < J; ++j) l[c][j] = v[j];
            for (int j = 0; j < J; ++j) {      // odometer, dim 0 fastest
                ++v[j];
                int sum = 0;
                for (int k = 0; k < J; ++k) sum += v[k];
                if (sum <= R) break;
                v[j] = 0;
            }
        }
    }
};

template <int J, int R, int c, int j>
__host__ __device__ constexpr int simplex_level() {
    return Simplex<J, R>().l[c][j];
}

// Lanes that set kPackedUb keep the offset table of tabled dimension 0 only; for the
// others they keep t_7 (the table's t at level 7) and take t_l = t_7 >> (7 - l): the
// table's recurrence t_{l+1} = 2 t_l + [m_l >= 0] makes the two equal.
template <class Ln, class = void>
struct PackedUb : std::false_type {};
template <class Ln>
struct PackedUb<Ln, std::void_t<decltype(Ln::kPackedUb)>>
    : std::integral_constant<bool, Ln::kPackedUb> {};

// Lanes that set kTopOnTheFly keep no tables for the outermost tabled dimension J - 1:
// its hat and offset at level l come from x at that level (hat_at), once per run of that
// level in a static block.
template <class Ln, class = void>
struct TopOnTheFly : std::false_type {};
template <class Ln>
struct TopOnTheFly<Ln, std::void_t<decltype(Ln::kTopOnTheFly)>>
    : std::integral_constant<bool, Ln::kTopOnTheFly> {};

template <class T>
__device__ __forceinline__ void hat_at(T x, int l, T& phi, uint32_t& t);

// whether tabled dimension j of the lane type takes its hats and offsets from x
template <class Ln, int j>
__host__ __device__ constexpr bool on_the_fly_dim() {
    if constexpr (TopOnTheFly<Ln>::value) return j == Ln::kJ - 1;
    else return false;
}

// byte offset t_l * sizeof(T) of tabled dimension j at level l
template <int j, int l, class Ln>
__device__ __forceinline__ uint32_t ub_at(const Ln& ln) {
    if constexpr (on_the_fly_dim<Ln, j>()) {
        std::remove_cv_t<std::remove_reference_t<decltype(ln.x[0])>> phi;
        uint32_t t;
        hat_at(ln.x[j], l, phi, t);
        return t * uint32_t(Ln::kElem);
    } else if constexpr (PackedUb<Ln>::value && j > 0) {
        return (ln.t7[j] >> (7 - l)) * uint32_t(Ln::kElem);
    } else {
        return ln.ub[j][l];
    }
}

// hat of tabled dimension j at level l
template <int j, int l, class Ln>
__device__ __forceinline__ auto ph_at(const Ln& ln) {
    if constexpr (on_the_fly_dim<Ln, j>()) {
        std::remove_cv_t<std::remove_reference_t<decltype(ln.x[0])>> phi;
        uint32_t t;
        hat_at(ln.x[j], l, phi, t);
        return phi;
    } else {
        return ln.ph[j][l];
    }
}

// the tables of tabled dimension k of a lane from its coordinate xk
template <class Ln, class T>
__device__ __forceinline__ void fill_tables(Ln& ln, int k, T xk) {
    if constexpr (PackedUb<Ln>::value) {
        if (k == 0) {
            level_table<8, T>(xk, ln.ph[0], ln.ub[0]);
        } else {
            uint32_t ub[8];
            level_table<8, T>(xk, ln.ph[k], ub);
            ln.t7[k] = ub[7] / uint32_t(sizeof(T));
        }
    } else {
        level_table<8, T>(xk, ln.ph[k], ln.ub[k]);
    }
}

// in-block byte offset of leaf c: dims J-1 .. 0, each shifted by the levels above it
template <int J, int R, int c, int j, class Ln>
__device__ __forceinline__ uint32_t leaf_offset(const Ln& ln, uint32_t off, int sh) {
    constexpr int l = simplex_level<J, R, c, j>();
    off += ub_at<j, l>(ln) << sh;
    if constexpr (j > 0) return leaf_offset<J, R, c, j - 1>(ln, off, sh + l);
    return off;
}

template <int J, int R, int c, int j, class Ln, class T>
__device__ __forceinline__ T leaf_weight(const Ln& ln, T w) {
    constexpr int l = simplex_level<J, R, c, j>();
    w = w * ph_at<j, l>(ln);
    if constexpr (j > 0) return leaf_weight<J, R, c, j - 1>(ln, w);
    return w;
}

template <int J, int R, int B, bool LDS, int c, int cend, class Ln, class T, class St>
__device__ __forceinline__ void block_issue(T (&a)[B], const Ln& ln, uint32_t off, int hsum,
                                            St& st) {
    if constexpr (c < cend) {
        const uint32_t o = leaf_offset<J, R, c, J - 1>(ln, off, hsum);
        if constexpr (LDS) a[c % B] = st.fetch_lds(st.q + c, o);
        else a[c % B] = st.fetch_global(st.q + c, o);
        block_issue<J, R, B, LDS, c + 1, cend>(a, ln, off, hsum, st);
    }
}

template <int J, int R, int B, int c, int cend, class Ln, class T, class Acc>
__device__ __forceinline__ void block_consume(const T (&a)[B], const Ln& ln, T pp, Acc& acc) {
    if constexpr (c < cend) {
        acc.add(leaf_weight<J, R, c, J - 1>(ln, pp), a[c % B]);
        block_consume<J, R, B, c + 1, cend>(a, ln, pp, acc);
    }
}

// FP64 blocks are summed nested instead of leaf by leaf:
//   pp * sum_{l_{J-1}} ph_{J-1} ( ... sum_{l_1} ph_1 ( sum_{l_0} ph_0 a ) ... ).
// s[0] runs over l_0 for fixed higher levels (consecutive leaves); when a run of level j
// ends its sum is folded into s[j+1] with weight ph_{j+1}; the block's sum times the
// prefix product pp goes into the accumulator.  One FMA per leaf and one per finished
// run, where the leaf-by-leaf form recomputes part of the hat product per leaf.
// Run boundaries are compile-time facts of the simplex enumeration.
template <int J, int R, int c, int j>
__host__ __device__ constexpr int run_max() {         // largest l_j given the levels above
    int m = R;
    for (int i = j + 1; i < J; ++i) m -= Simplex<J, R>().l[c][i];
    return m;
}

template <int J, int R, int c, int j, class Ln, class T, class Acc>
__device__ __forceinline__ void nested_fold(T (&sn)[J], const Ln& ln, T pp, Acc& acc) {
    if constexpr (j == J) {
        acc.s = fma(pp, sn[J - 1], acc.s);
    } else {
        constexpr int l = simplex_level<J, R, c, j>();
        if constexpr (l == 0) sn[j] = ph_at<j, 0>(ln) * sn[j - 1];
        else sn[j] = fma(ph_at<j, l>(ln), sn[j - 1], sn[j]);
        if constexpr (l == run_max<J, R, c, j>()) nested_fold<J, R, c, j + 1>(sn, ln, pp, acc);
    }
}

template <int J, int R, int B, int c, int cend, class Ln, class T, class Acc>
__device__ __forceinline__ void block_consume_nested(const T (&a)[B], T (&sn)[J], const Ln& ln,
                                                     T pp, Acc& acc) {
    if constexpr (c < cend) {
        constexpr int l0 = simplex_level<J, R, c, 0>();
        if constexpr (l0 == 0) sn[0] = ph_at<0, 0>(ln) * a[c % B];
        else sn[0] = fma(ph_at<0, l0>(ln), a[c % B], sn[0]);
        if constexpr (l0 == run_max<J, R, c, 0>()) nested_fold<J, R, c, 1>(sn, ln, pp, acc);
        block_consume_nested<J, R, B, c + 1, cend>(a, sn, ln, pp, acc);
    }
}

template <int J, int R, int B, bool LDS, int c0, class Ln, class T, class St>
__device__ __forceinline__ void block_batches(const Ln& ln, T pp, uint32_t off, int hsum,
                                              St& st, T (&sn)[J]) {
    constexpr int C = Simplex<J, R>::C;
    if constexpr (c0 < C) {
        constexpr int cend = (c0 + B < C) ? c0 + B : C;
        T a[B];
        block_issue<J, R, B, LDS, c0, cend>(a, ln, off, hsum, st);
        if constexpr (kNestedSum<decltype(st.acc)>)
            block_consume_nested<J, R, B, c0, cend>(a, sn, ln, pp, st.acc);
        else
            block_consume<J, R, B, c0, cend>(a, ln, pp, st.acc);
        block_batches<J, R, B, LDS, cend>(ln, pp, off, hsum, st, sn);
    }
}

// Block-staged static blocks: the outer level sum H is a compile-time constant, so every
// shift of the in-block offset is one too, and the leaf's place in the block's LDS image
// (blk_leaf_offset) becomes the LDS read's immediate offset.
template <int J, int R, int c, int j, int SH, class Ln>
__device__ __forceinline__ uint32_t leaf_offset_ct(const Ln& ln, uint32_t off) {
    constexpr int l = simplex_level<J, R, c, j>();
    off += ub_at<j, l>(ln) << SH;
    if constexpr (j > 0) return leaf_offset_ct<J, R, c, j - 1, SH + l>(ln, off);
    return off;
}

template <int J, int R, int c>
__host__ __device__ constexpr int64_t blk_leaf_lds(int H, int elem) {
    int lv[J] = {};
    for (int j = 0; j < J; ++j) lv[j] = Simplex<J, R>().l[c][j];
    return blk_leaf_offset(J, lv, H, elem);
}

template <int J, int R, int H, int B, int c, int cend, class Ln, class T, class St>
__device__ __forceinline__ void block_issue_blk(T (&a)[B], const Ln& ln, uint32_t vb, St& st) {
    if constexpr (c < cend) {
        constexpr uint32_t lo = uint32_t(blk_leaf_lds<J, R, c>(H, int(sizeof(T))));
        a[c % B] = st.fetch_lds_at(leaf_offset_ct<J, R, c, J - 1, H>(ln, vb) + lo);
        block_issue_blk<J, R, H, B, c + 1, cend>(a, ln, vb, st);
    }
}

// With H a compile-time constant a leaf's address minus the block base depends on the
// lane's tables alone, so the compiler would hoist the offsets of every leaf of every
// block out of the walk and keep them all live.  An empty asm that claims to change the
// offset tables at each block entry keeps that arithmetic inside the block.
template <int J, int R, class Ln>
__device__ __forceinline__ void launder_offsets(Ln& ln) {
    if constexpr (TopOnTheFly<Ln>::value) asm volatile("" : "+v"(ln.x[J - 1]));
    if constexpr (PackedUb<Ln>::value) {
#pragma unroll
        for (int l = 0; l <= R && l < 8; ++l) asm volatile("" : "+v"(ln.ub[0][l]));
#pragma unroll
        for (int j = 1; j < J - (TopOnTheFly<Ln>::value ? 1 : 0); ++j)
            asm volatile("" : "+v"(ln.t7[j]));
    } else {
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int l = 0; l <= R && l < int(sizeof(ln.ub[0]) / sizeof(ln.ub[0][0])); ++l)
                asm volatile("" : "+v"(ln.ub[j][l]));
    }
}

template <int J, int R, int H, int B, int c0, class Ln, class T, class St>
__device__ __forceinline__ void block_batches_blk(const Ln& ln, T pp, uint32_t vb, St& st,
                                                  T (&sn)[J]) {
    constexpr int C = Simplex<J, R>::C;
    if constexpr (c0 < C) {
        constexpr int cend = (c0 + B < C) ? c0 + B : C;
        T a[B];
        block_issue_blk<J, R, H, B, c0, cend>(a, ln, vb, st);
        if constexpr (kNestedSum<decltype(st.acc)>)
            block_consume_nested<J, R, B, c0, cend>(a, sn, ln, pp, st.acc);
        else
            block_consume<J, R, B, c0, cend>(a, ln, pp, st.acc);
        block_batches_blk<J, R, H, B, cend>(ln, pp, vb, st, sn);
    }
}

// A block too large for a stage is staged by slices: slice a holds the leaves with
// l_{J-1} = a, laid out as a block over the J-1 inner tabled dimensions below outer level
// sum H + a, and every slice takes its own place in the stages (possibly the next
// stage).  The leaves, their order and the block's arithmetic (nested sums carried from
// slice to slice) are those of the whole block; only where the reads come from changes.
template <int J, int R>
__host__ __device__ constexpr int slice_first(int a) {     // first leaf with l_{J-1} = a
    int c = 0;
    for (int b = 0; b < a; ++b) c += binom(R - b + J - 1, J - 1);
    return c;
}

template <int J, int R, int c>
__host__ __device__ constexpr int64_t slice_leaf_lds(int H, int elem) {
    int lv[J] = {};
    for (int j = 0; j < J - 1; ++j) lv[j] = Simplex<J, R>().l[c][j];
    return blk_leaf_offset(J - 1, lv, H + Simplex<J, R>().l[c][J - 1], elem);
}

template <int J, int R, int H, int B, int c, int cend, class Ln, class T, class St>
__device__ __forceinline__ void slice_issue(T (&a)[B], const Ln& ln, uint32_t vb, St& st) {
    if constexpr (c < cend) {
        constexpr uint32_t lo = uint32_t(slice_leaf_lds<J, R, c>(H, int(sizeof(T))));
        a[c % B] = st.fetch_lds_at(leaf_offset_ct<J, R, c, J - 1, H>(ln, vb) + lo);
        slice_issue<J, R, H, B, c + 1, cend>(a, ln, vb, st);
    }
}

template <int J, int R, int H, int B, int c0, int c1, class Ln, class T, class St>
__device__ __forceinline__ void slice_batches(const Ln& ln, T pp, uint32_t vb, St& st,
                                              T (&sn)[J]) {
    if constexpr (c0 < c1) {
        constexpr int cend = (c0 + B < c1) ? c0 + B : c1;
        T a[B];
        slice_issue<J, R, H, B, c0, cend>(a, ln, vb, st);
        if constexpr (kNestedSum<decltype(st.acc)>)
            block_consume_nested<J, R, B, c0, cend>(a, sn, ln, pp, st.acc);
        else
            block_consume<J, R, B, c0, cend>(a, ln, pp, st.acc);
        // a slice is one long straight-line stretch: keep each batch's reads with it (the
        // scheduler otherwise hoists the whole slice's reads and the registers spill)
        __builtin_amdgcn_sched_barrier(0);
        slice_batches<J, R, H, B, cend, c1>(ln, pp, vb, st, sn);
    }
}

template <int J, int R, int H, int B, int a, class Ln, class T, class St>
__device__ __forceinline__ void block_slices(const Ln& ln, T pp, uint32_t off, St& st,
                                             T (&sn)[J]) {
    if constexpr (a <= R) {
        // the nested sums of the slices so far must exist here, before the stage check: the
        // compiler could otherwise sink their FMAs past it to the block's end and keep all
        // the slices' loaded values alive
#pragma unroll
        for (int j = 0; j < J; ++j) asm volatile("" : "+v"(sn[j]));
        st.template block_begin<uint32_t(blk_bytes(J - 1, R - a, H + a, int(sizeof(T))))>();
        launder_offsets<J, R>(const_cast<Ln&>(ln));
        slice_batches<J, R, H, B, slice_first<J, R>(a), slice_first<J, R>(a + 1)>(
            ln, pp, off + st.base, st, sn);
        block_slices<J, R, H, B, a + 1>(ln, pp, off, st, sn);
    }
}

// All C(R+J, J) leaves below a depth-J node whose remaining level budget R fits inside
// the J tabled dimensions.  The staged walk reads a block from LDS when the host packed
// it into a stage (the same rule, stage_whole, decides on both sides), else from global
// memory.
template <int J, int R, int B, class Ln, class T, class St>
__device__ __forceinline__ void static_block(const Ln& ln, T pp, uint32_t off, int hsum,
                                             St& st) {
    T sn[J];                   // nested partial sums (FP64 blocks)
    if constexpr (St::kBlk) {
        constexpr int H = St::kNM1 - R;
        constexpr int kHow = blk_staging(J, R, H, int(sizeof(T)), St::kStageBytes);
        if constexpr (kHow == 1) {
            st.template block_begin<uint32_t(blk_bytes(J, R, H, int(sizeof(T))))>();
            launder_offsets<J, R>(const_cast<Ln&>(ln));   // ln is the kernel's own lane
            block_batches_blk<J, R, H, B, 0>(ln, pp, off + st.base, st, sn);
        } else if constexpr (kHow == 2) {
            block_slices<J, R, H, B, 0>(ln, pp, off, st, sn);
        } else {
            block_batches<J, R, B, false, 0>(ln, pp, off, hsum, st, sn);
        }
    } else if constexpr (St::kStaged) {
        constexpr int kHmax = stage_hmax(J, R, int(sizeof(T)));
        if (hsum <= kHmax) {
            st.block_start(st.q);
            block_batches<J, R, B, true, 0>(ln, pp, off, hsum, st, sn);
        } else {
            block_batches<J, R, B, false, 0>(ln, pp, off, hsum, st, sn);
        }
    } else {
        block_batches<J, R, B, false, 0>(ln, pp, off, hsum, st, sn);
    }
    st.q += Simplex<J, R>::C;
}

template <int J, int LT, int B, class Ln, class T, class St>
__device__ __forceinline__ void static_block_dispatch(int R, const Ln& ln, T pp, uint32_t off,
                                                      int hsum, St& st) {
    switch (R) {
#define SG_BLK(RR) \
    case RR: if constexpr (RR < LT) static_block<J, RR, B>(ln, pp, off, hsum, st); break;
        SG_BLK(0) SG_BLK(1) SG_BLK(2) SG_BLK(3) SG_BLK(4) SG_BLK(5) SG_BLK(6) SG_BLK(7)
#undef SG_BLK
        default: break;
    }
}

// The leaves below a depth-(J+1) node whose budget R (at most up_max) fits dimension J
// and the J tabled ones: dimension J's levels 0..R unrolled at compile time, its hat
// advanced incrementally as in the walk, each level a static block of budget R - l.
// Saves the walk's per-node dispatch for the many small blocks; leaves, their order and
// their arithmetic are those of the walk.
//   FP32 walks unroll budgets up to 4 (4-26 % faster at d = 5..12).
//   FP64 registers are tighter: every unrolled budget adds spill traffic, so only the
//   staged walk unrolls, and less as d grows (it reads in batches of 4 from d = 9 on,
//   which frees registers for it).  Measured on regular grids, 2^22-2^24 points: budget
//   3 up to d = 8, 4 at d = 9, 3 at d = 10 and 11 (d = 10, n = 8: 182 ms unrolled
//   nowhere, 147 ms; with budget 4 there, 149 ms), 2 at d = 12..14, 1 at d = 15 and 16,
//   none at d = 20 (a budget of 1 there is 16 % slower).  The gathering FP64 variants
//   do not unroll (up to 2x slower at d = 16).
template <int D, class T, bool STG>
constexpr int up_max() {
    if constexpr (sizeof(T) == 4) return 4;
    else if constexpr (STG)
        return D <= 8 ? 3 : D == 9 ? 4 : D <= 11 ? 3 : D <= 14 ? 2 : D <= 16 ? 1 : 0;
    else return 0;
}

// Batch of the staged walk (leaves whose reads are issued before the first is used).
// Most of its reads come from LDS and need few in flight; from d = 9 on, FP64 batches of
// 4 free registers for the unrolled levels (d = 12..16: 9-21 % faster than 8; d = 6..8:
// 3 % slower).
template <int D, class T>
constexpr int stage_batch() {
    return sizeof(T) == 8 && D >= 9 ? 4 : 8;
}
// batch of the block-staged walk
#ifndef SG_BLK_BATCH64
#define SG_BLK_BATCH64 4
#endif
#ifndef SG_BLK_BATCH32
#define SG_BLK_BATCH32 8
#endif
template <int D, class T>
constexpr int blk_batch() {
    return sizeof(T) == 8 ? SG_BLK_BATCH64 : SG_BLK_BATCH32;
}

// Up to two dimensions above the tabled ones unroll this way: dimension K's levels
// 0..R, each level unrolling dimension K-1 with the budget left, down to dimension J
// whose levels each end in a static block.
template <int K, int J, int R, int B, int l, class Ln, class T, class St>
__device__ __forceinline__ void up_block_level(const Ln& ln, T pp, uint32_t off, int hsum, T m,
                                               uint32_t t, St& st) {
    if constexpr (l <= R) {
        const T phi = T(1) - fabs(m);
        const uint32_t o = off + ((t * uint32_t(sizeof(T))) << hsum);
        if constexpr (K == J)
            static_block<J, R - l, B>(ln, pp * phi, o, hsum + l, st);
        else
            up_block_level<K - 1, J, R - l, B, 0>(ln, pp * phi, o, hsum + l, ln.m0[K - 1], 0u, st);
        const bool ge = m >= T(0);
        up_block_level<K, J, R, B, l + 1>(ln, pp, off, hsum, T(2) * m + (ge ? T(-1) : T(1)),
                                          2u * t + (ge ? 1u : 0u), st);
    }
}

template <int K, int J, int LT, int B, int UP, class Ln, class T, class St>
__device__ __forceinline__ void up_block_dispatch(int R, const Ln& ln, T pp, uint32_t off,
                                                  int hsum, St& st) {
    switch (R) {
#define SG_UP(RR) \
    case RR: if constexpr (RR < LT && RR <= UP) up_block_level<K, J, RR, B, 0>(ln, pp, off, hsum, ln.m0[K], 0u, st); break;
        SG_UP(0) SG_UP(1) SG_UP(2) SG_UP(3) SG_UP(4)
#undef SG_UP
        default: break;
    }
}

// Largest budget for which dimensions J+1 and J both unroll (0: never); staged walks
// only.  Measured on regular grids (ms, 2^24 points at d = 10, 2^22 otherwise):
//   FP64, budget 0 / 1 / 2: d = 10, n = 8: 150.7 / 153.2 / 148.7; d = 11, n = 7: 26.7 /
//   25.7 / 26.7; d = 12, n = 7: 45.6 / 44.9 / 43.7; d = 13, n = 6: 28.0 / 27.5 / 30.5;
//   elsewhere the extra live registers cost up to 1.6x (d = 20).
//   FP32 (Kahan), budget 0 / 2 / 4: d = 10: 124.1 / 111.0 / 108.5; d = 12: 38.8 / 32.4 /
//   31.5; d = 14, n = 6: 36.7 / 28.9 / 28.7; d = 16, n = 5: 26.1 / 20.2 / 20.5.
template <int D, class T, bool STG>
constexpr int up2_max() {
    if constexpr (!STG) return 0;
    else if constexpr (sizeof(T) == 4) return 4;
    else return D == 10 || D == 12 ? 2 : D == 11 || D == 13 ? 1 : 0;
}

// Largest budget for which the general walk unrolls dimensions J+1 and J together when it
// runs the block-staged blocks (FP32): 2 (the staged walk's 4 gains under 2 % over 2 and
// doubles the kernels' code and build time).
#ifndef SG_BLK_UP2_32
#define SG_BLK_UP2_32 2
#endif
template <class T>
constexpr int blk_up2_max() {
    return sizeof(T) == 4 ? SG_BLK_UP2_32 : 0;
}

// The block-staged walk of a regular grid (every limit n - 1): every node below the
// root admits levels 0..rem, and a depth-J node is always a whole static block.  Each
// outer hat is computed from x at its level (y = x 2^l, t = floor(y), hat = 1 - |2(y - t)
// - 1|) instead of incrementally, so no level loop carries hat state across the walk
// below it: the live state per open level is its prefix product and offset.  For x on
// the 2^-24 lattice both forms are exact, so the leaves' weights equal the other walks'.
template <class T>
__device__ __forceinline__ void hat_at(T x, int l, T& phi, uint32_t& t) {
    const T y = ldexp(x, l);
    const T fl = floor(y);
    t = uint32_t(fl);
    phi = T(1) - fabs(fma(T(2), y - fl, T(-1)));
}

// Lane state of the block-staged walk: the coordinates and the J tabled dimensions'
// hat / offset tables.
// PACK (the walks at 4 waves per SIMD, blk_quad): the offset table of dimension 0 only,
// t_7 of the other tabled dimensions (ub_at), and no tables for the outermost tabled
// dimension, whose hats and offsets come from x once per run (ph_at, ub_at): 162 -> 128
// VGPRs at d = 10, n = 8.  At 3 waves per SIMD the recomputed hats cost 2 % (146.8 ms
// against 144.1).
template <int D, int J, class T, bool PACK = false>
struct LaneBlk {
    static constexpr bool kPackedUb = PACK;
    static constexpr bool kTopOnTheFly = PACK && J > 1;
    static constexpr int kJ = J;
    static constexpr int kElem = int(sizeof(T));
    T x[D];
    T ph[J][8];
    uint32_t ub[PACK ? 1 : J][8];
    uint32_t t7[J];             // PACK: t_7 of each tabled dimension
};

// Which walk the block-staged kernels use.  FP64 takes the lean regular walk above
// (hats computed at each level from x, nothing carried across a level loop: 158 VGPRs at
// d = 10, no spills; unrolling any level above the blocks made the compiler spill).  FP32
// has registers to spare and takes the general walk with its unrolled levels and
// incremental hats (d = 10, n = 8, 2^24 points: 95.8 ms against 133 ms for the lean walk).
#ifndef SG_BLK_LEAN64
#define SG_BLK_LEAN64 1
#endif
template <class T>
constexpr bool blk_lean_walk() {
    return sizeof(T) == 8 && SG_BLK_LEAN64;
}
// Dimension J's levels 0..rem above the static blocks, as one switch that falls through:
// entering at case rem runs the static blocks of budgets rem, rem - 1, ..., 0 (level
// l = rem - R of dimension J each), so a node of dimension J+1 dispatches once instead
// of once per static block, and every static block's code exists once.  Used up to
// d = 10 (d = 10, n = 8: 146.6 -> 144.1 ms); from d = 12 on it is neutral to 15 % slower.
#ifndef SG_BLK_FALL_DMAX
#define SG_BLK_FALL_DMAX 10
#endif
template <int J, int B, int R, class Ln, class T, class St>
__device__ __forceinline__ void blk_level_run(Ln& ln, T pp, uint32_t off, int hsum, int rem,
                                              St& st) {
    T phi;
    uint32_t t;
    const int l = rem - R;
    hat_at(ln.x[J], l, phi, t);
    const uint32_t o = off + ((t * uint32_t(sizeof(T))) << hsum);
    static_block<J, R, B>(ln, pp * phi, o, hsum + l, st);
}

template <int J, int B, class Ln, class T, class St>
__device__ __forceinline__ void blk_levels(Ln& ln, T pp, uint32_t off, int hsum, int rem,
                                           St& st) {
    switch (rem) {
#define SG_LV(RR) \
    case RR: blk_level_run<J, B, RR>(ln, pp, off, hsum, rem, st); [[fallthrough]];
        SG_LV(7) SG_LV(6) SG_LV(5) SG_LV(4) SG_LV(3) SG_LV(2) SG_LV(1) SG_LV(0)
#undef SG_LV
        default: break;
    }
}

template <int K, int D, int J, int B, class T, class St, class Ln = LaneBlk<D, J, T>>
struct WalkerBlk {
    static __device__ __forceinline__ void run(Ln& ln, T pp, uint32_t off, int hsum, int rem,
                                               St& st) {
        if constexpr (K == J - 1) {
            static_block_dispatch<J, 8, B>(rem, ln, pp, off, hsum, st);
        } else if constexpr (K == J && D <= SG_BLK_FALL_DMAX) {
            blk_levels<J, B>(ln, pp, off, hsum, rem, st);
        } else {
            for (int l = 0; l <= rem; ++l) {
                T phi;
                uint32_t t;
                hat_at(ln.x[K], l, phi, t);
                const uint32_t o = off + ((t * uint32_t(sizeof(T))) << hsum);
                WalkerBlk<K - 1, D, J, B, T, St, Ln>::run(ln, pp * phi, o, hsum + l, rem - l,
                                                          st);
            }
        }
    }
};

// ------------------------------------------------------ two points per lane --
// The block-staged walk with two query points per lane (variant 5): a workgroup of 256
// lanes evaluates 512 points.  Both points of a lane visit the same leaves in the same
// order, so the stage copies, the stage barriers and the walk's scalar work (loop control,
// stage checks) are shared by twice the points, while every per-point operation (hats,
// addresses, reads, sums) is the single-point walk's, done once per point.  Each point's
// arithmetic is exactly the lean walk's, so the outputs are bit-identical to variant 4's
// (and variant 0's).

// the two points' accumulators; block_begin pins both
template <class Acc>
struct AccPair {
    using Inner = Acc;
    Acc a[2];
    __device__ __forceinline__ void pin() {
        a[0].pin();
        a[1].pin();
    }
};

template <int D, int J, class T>
struct LaneBlk2 {
    LaneBlk<D, J, T> p[2];
};

template <int J, int R, int B, int c0, int cend, class Ln2, class T, class Acc>
__device__ __forceinline__ void consume2(const T (&a0)[B], const T (&a1)[B], T (&sn0)[J],
                                         T (&sn1)[J], const Ln2& ln, T pp0, T pp1,
                                         AccPair<Acc>& acc) {
    if constexpr (kNestedSum<Acc>) {
        block_consume_nested<J, R, B, c0, cend>(a0, sn0, ln.p[0], pp0, acc.a[0]);
        block_consume_nested<J, R, B, c0, cend>(a1, sn1, ln.p[1], pp1, acc.a[1]);
    } else {
        block_consume<J, R, B, c0, cend>(a0, ln.p[0], pp0, acc.a[0]);
        block_consume<J, R, B, c0, cend>(a1, ln.p[1], pp1, acc.a[1]);
    }
}

// a whole staged block at LDS offsets vb0 / vb1 (block base + each point's in-block offset)
template <int J, int R, int H, int B, int c0, class Ln2, class T, class St>
__device__ __forceinline__ void block_batches_blk2(const Ln2& ln, T pp0, T pp1, uint32_t vb0,
                                                   uint32_t vb1, St& st, T (&sn0)[J],
                                                   T (&sn1)[J]) {
    constexpr int C = Simplex<J, R>::C;
    if constexpr (c0 < C) {
        constexpr int cend = (c0 + B < C) ? c0 + B : C;
        T a0[B], a1[B];
        block_issue_blk<J, R, H, B, c0, cend>(a0, ln.p[0], vb0, st);
        block_issue_blk<J, R, H, B, c0, cend>(a1, ln.p[1], vb1, st);
        consume2<J, R, B, c0, cend>(a0, a1, sn0, sn1, ln, pp0, pp1, st.acc);
        block_batches_blk2<J, R, H, B, cend>(ln, pp0, pp1, vb0, vb1, st, sn0, sn1);
    }
}

template <int J, int R, int H, int B, int c0, int c1, class Ln2, class T, class St>
__device__ __forceinline__ void slice_batches2(const Ln2& ln, T pp0, T pp1, uint32_t vb0,
                                               uint32_t vb1, St& st, T (&sn0)[J], T (&sn1)[J]) {
    if constexpr (c0 < c1) {
        constexpr int cend = (c0 + B < c1) ? c0 + B : c1;
        T a0[B], a1[B];
        slice_issue<J, R, H, B, c0, cend>(a0, ln.p[0], vb0, st);
        slice_issue<J, R, H, B, c0, cend>(a1, ln.p[1], vb1, st);
        consume2<J, R, B, c0, cend>(a0, a1, sn0, sn1, ln, pp0, pp1, st.acc);
        __builtin_amdgcn_sched_barrier(0);
        slice_batches2<J, R, H, B, cend, c1>(ln, pp0, pp1, vb0, vb1, st, sn0, sn1);
    }
}

// a block too large for a stage, slice by slice (block_slices)
template <int J, int R, int H, int B, int a, class Ln2, class T, class St>
__device__ __forceinline__ void block_slices2(const Ln2& ln, T pp0, T pp1, uint32_t off0,
                                              uint32_t off1, St& st, T (&sn0)[J], T (&sn1)[J]) {
    if constexpr (a <= R) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            asm volatile("" : "+v"(sn0[j]));
            asm volatile("" : "+v"(sn1[j]));
        }
        st.template block_begin<uint32_t(blk_bytes(J - 1, R - a, H + a, int(sizeof(T))))>();
        launder_offsets<J, R>(const_cast<Ln2&>(ln).p[0]);
        launder_offsets<J, R>(const_cast<Ln2&>(ln).p[1]);
        slice_batches2<J, R, H, B, slice_first<J, R>(a), slice_first<J, R>(a + 1)>(
            ln, pp0, pp1, off0 + st.base, off1 + st.base, st, sn0, sn1);
        block_slices2<J, R, H, B, a + 1>(ln, pp0, pp1, off0, off1, st, sn0, sn1);
    }
}

// a block gathered from global memory
template <int J, int R, int B, int c0, class Ln2, class T, class St>
__device__ __forceinline__ void block_batches2_global(const Ln2& ln, T pp0, T pp1, uint32_t off0,
                                                      uint32_t off1, int hsum, St& st,
                                                      T (&sn0)[J], T (&sn1)[J]) {
    constexpr int C = Simplex<J, R>::C;
    if constexpr (c0 < C) {
        constexpr int cend = (c0 + B < C) ? c0 + B : C;
        T a0[B], a1[B];
        block_issue<J, R, B, false, c0, cend>(a0, ln.p[0], off0, hsum, st);
        block_issue<J, R, B, false, c0, cend>(a1, ln.p[1], off1, hsum, st);
        consume2<J, R, B, c0, cend>(a0, a1, sn0, sn1, ln, pp0, pp1, st.acc);
        block_batches2_global<J, R, B, cend>(ln, pp0, pp1, off0, off1, hsum, st, sn0, sn1);
    }
}

template <int J, int R, int B, class Ln2, class T, class St>
__device__ __forceinline__ void static_block2(const Ln2& ln, T pp0, T pp1, uint32_t off0,
                                              uint32_t off1, int hsum, St& st) {
    static_assert(St::kBlk, "two points per lane run the block-staged walk");
    constexpr int H = St::kNM1 - R;
    constexpr int E = int(sizeof(T));
    constexpr int kHow = blk_staging(J, R, H, E, St::kStageBytes);
    T sn0[J], sn1[J];
    if constexpr (kHow == 1) {
        st.template block_begin<uint32_t(blk_bytes(J, R, H, E))>();
        launder_offsets<J, R>(const_cast<Ln2&>(ln).p[0]);
        launder_offsets<J, R>(const_cast<Ln2&>(ln).p[1]);
        block_batches_blk2<J, R, H, B, 0>(ln, pp0, pp1, off0 + st.base, off1 + st.base, st, sn0,
                                          sn1);
    } else if constexpr (kHow == 2) {
        block_slices2<J, R, H, B, 0>(ln, pp0, pp1, off0, off1, st, sn0, sn1);
    } else {
        block_batches2_global<J, R, B, 0>(ln, pp0, pp1, off0, off1, hsum, st, sn0, sn1);
    }
    st.q += Simplex<J, R>::C;
}

template <int J, int LT, int B, class Ln2, class T, class St>
__device__ __forceinline__ void static_block_dispatch2(int R, const Ln2& ln, T pp0, T pp1,
                                                       uint32_t off0, uint32_t off1, int hsum,
                                                       St& st) {
    switch (R) {
#define SG_BLK2(RR) \
    case RR: if constexpr (RR < LT) static_block2<J, RR, B>(ln, pp0, pp1, off0, off1, hsum, st); break;
        SG_BLK2(0) SG_BLK2(1) SG_BLK2(2) SG_BLK2(3) SG_BLK2(4) SG_BLK2(5) SG_BLK2(6) SG_BLK2(7)
#undef SG_BLK2
        default: break;
    }
}

template <int J, int B, int R, class Ln2, class T, class St>
__device__ __forceinline__ void blk_level_run2(Ln2& ln, T pp0, T pp1, uint32_t off0,
                                               uint32_t off1, int hsum, int rem, St& st) {
    T phi0, phi1;
    uint32_t t0, t1;
    const int l = rem - R;
    hat_at(ln.p[0].x[J], l, phi0, t0);
    hat_at(ln.p[1].x[J], l, phi1, t1);
    static_block2<J, R, B>(ln, pp0 * phi0, pp1 * phi1,
                           off0 + ((t0 * uint32_t(sizeof(T))) << hsum),
                           off1 + ((t1 * uint32_t(sizeof(T))) << hsum), hsum + l, st);
}

template <int J, int B, class Ln2, class T, class St>
__device__ __forceinline__ void blk_levels2(Ln2& ln, T pp0, T pp1, uint32_t off0, uint32_t off1,
                                            int hsum, int rem, St& st) {
    switch (rem) {
#define SG_LV2(RR) \
    case RR: blk_level_run2<J, B, RR>(ln, pp0, pp1, off0, off1, hsum, rem, st); [[fallthrough]];
        SG_LV2(7) SG_LV2(6) SG_LV2(5) SG_LV2(4) SG_LV2(3) SG_LV2(2) SG_LV2(1) SG_LV2(0)
#undef SG_LV2
        default: break;
    }
}

// the lean block-staged walk (WalkerBlk) for both points of a lane
template <int K, int D, int J, int B, class T, class St>
struct WalkerBlk2 {
    using Ln2 = LaneBlk2<D, J, T>;
    static __device__ __forceinline__ void run(Ln2& ln, T pp0, T pp1, uint32_t off0,
                                               uint32_t off1, int hsum, int rem, St& st) {
        if constexpr (K == J - 1) {
            static_block_dispatch2<J, 8, B>(rem, ln, pp0, pp1, off0, off1, hsum, st);
        } else if constexpr (K == J && D <= SG_BLK_FALL_DMAX) {
            blk_levels2<J, B>(ln, pp0, pp1, off0, off1, hsum, rem, st);
        } else {
            for (int l = 0; l <= rem; ++l) {
                T phi0, phi1;
                uint32_t t0, t1;
                hat_at(ln.p[0].x[K], l, phi0, t0);
                hat_at(ln.p[1].x[K], l, phi1, t1);
                WalkerBlk2<K - 1, D, J, B, T, St>::run(
                    ln, pp0 * phi0, pp1 * phi1, off0 + ((t0 * uint32_t(sizeof(T))) << hsum),
                    off1 + ((t1 * uint32_t(sizeof(T))) << hsum), hsum + l, rem - l, st);
            }
        }
    }
};

// Lane state of the sparse block-staged walk: the J tabled dimensions' hat / offset
// tables.  The coordinates of the outer dimensions J..D-1, which the walk indexes with a
// wave-uniform dimension known only at run time, live in separate kernel-local arrays of
// 32-bit words (FP64: low and high halves): those the compiler indexes in registers
// (s_set_gpr_idx), where an array of doubles, or a struct member, indexed at run time
// would keep its whole object in scratch memory.
// PACK: the packed tables of LaneBlk (ub_at, ph_at), which the sparse walk always takes:
// its nested levels spill less (d = 20, n = 6: 240 -> 140 bytes per lane; d = 16, n = 8:
// 347 -> 178 ms per 2^20 FP64 points; d = 12, n = 7: 43.4 -> 40.0 ms per 2^22; FP32
// d = 20, n = 6: 144.3 -> 141.2 ms; d = 20, n = 6 FP64 and d = 16, n = 5 unchanged).
template <int J, class T, bool PACK = true>
struct LaneSp {
    static constexpr bool kPackedUb = PACK;
    static constexpr bool kTopOnTheFly = PACK && J > 1;
    static constexpr int kJ = J;
    static constexpr int kElem = int(sizeof(T));
    T ph[J][8];
    uint32_t ub[PACK ? 1 : J][8];
    uint32_t t7[J];             // PACK: t_7 of each tabled dimension
    T x[J];                     // PACK: the outermost tabled dimension's coordinate
};
template <class T>
__device__ __forceinline__ T coord_at(const uint32_t* lo, const uint32_t* hi, int i) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, (uint64_t(hi[i]) << 32) | lo[i]);
    else
        return __builtin_bit_cast(float, lo[i]);
}

// The sparse order of the block-staged walk (blk_sparse; the host's sparse_walk lists the
// same blocks in the same order).  A node at nesting depth DEPTH has its dimensions >= c
// fixed, with P the product of their hats (DEPTH of them nonzero, so its budget rem is
// at most NM1 - DEPTH).  For k = c-1 down to J it visits the children with l_k = 1..rem,
// each with weight W * hat_k(l_k), where W = P * hat_{c-1}(0) * ... * hat_{k+1}(0) grows by
// one level-0 hat per k (also when rem = 0: the node's block needs all of them); then its
// own static block (dimensions J..c-1 at level 0) with the full product W.  Each weight is the product of the same hats, in the same
// order (outer dimensions descending), as in the depth-first walk.  Level 0 adds nothing
// to the in-block offset or the level sum, so dimension k's index bits shift by hsum, the
// sum of the nonzero levels above it.
// A node with no budget left (most nodes at high d: 20,349 of 26,334 at d = 20, n = 6)
// has a one-leaf block with every dimension 0..c-1 at level 0; its weight takes the
// per-point product Z_c = hat_0(0) * ... * hat_{c-1}(0) from a table instead of c - J
// steps of the loop and J multiplies in the block.  That product's rounding differs from
// the loop's (so from the depth-first walk's), well inside the FP64 bound.
template <int J, int NM1, int B, class Ln, class T, class St>
__device__ __forceinline__ void single_leaf(Ln& ln, T w, uint32_t off, int hsum, St& st) {
    constexpr int E = int(sizeof(T));
    if constexpr (blk_staging(J, 0, NM1, E, St::kStageBytes) == 1) {
        // the leaf (0, ..., 0) sits at the start of its block's LDS image
        st.template block_begin<uint32_t(blk_bytes(J, 0, NM1, E))>();
        st.acc.add(w, st.fetch_lds_at(off + st.base));
        st.q += 1;
    } else {
        st.acc.add(w, st.fetch_global(st.q, off));
        st.q += 1;
    }
}

template <int DEPTH, int NM1, int J, int B, class T, class St, class Ln>
struct WalkerSp {
    // xlo / xhi: coordinates of dimensions J.., zlo / zhi: Z_J..Z_D (coord_at); rem >= 1
    static __device__ __forceinline__ void run(Ln& ln, const uint32_t* xlo, const uint32_t* xhi,
                                               const uint32_t* zlo, const uint32_t* zhi, int c,
                                               T P, uint32_t off, int hsum, int rem, St& st) {
        T W = P;
        for (int k = c - 1; k >= J; --k) {
            const T xk = coord_at<T>(xlo, xhi, k - J);
            for (int l = 1; l <= rem; ++l) {
                T phi;
                uint32_t t;
                hat_at(xk, l, phi, t);
                const uint32_t o = off + ((t * uint32_t(sizeof(T))) << hsum);
                if (l == rem) {
                    // the child has no budget left: its one-leaf block, every tabled
                    // dimension at level 0 (in-block offset 0, hats folded into Z)
                    single_leaf<J, NM1, B>(ln, (W * phi) * coord_at<T>(zlo, zhi, k - J), o,
                                           hsum + l, st);
                } else {
                    if constexpr (DEPTH + 1 < NM1)
                        WalkerSp<DEPTH + 1, NM1, J, B, T, St, Ln>::run(ln, xlo, xhi, zlo, zhi, k,
                                                                       W * phi, o, hsum + l,
                                                                       rem - l, st);
                }
            }
            T phi0;
            uint32_t t0;
            hat_at(xk, 0, phi0, t0);
            W = W * phi0;
        }
        static_block_dispatch<J, NM1 - DEPTH + 1, B>(rem, ln, W, off, hsum, st);
    }
};

// Depth-first walk.  Dimension K's level runs 0..bound, where bound = min(rem,
// limit_K) for a restricted-regular set and comes from the node-code stream for an
// explicit set.  Dimension d-1 is outermost, so for a fixed level sum the leaves come
// in rank order.  pp is the product of the hats of dims > K, off the in-block byte
// offset they contribute, hsum the sum of their levels (the shift of dimension K's
// index bits), rem the level-sum budget left.
//
// Node codes of an explicit set (one byte per visited node, walk order): at depth J-1
// a code below 0x80 is a full static block of budget R = code; otherwise code & 0x7f is
// the largest l_K under this node.
template <int K, int D, int LT, int J, int B, bool EXPL, class T, class St,
          class Ln = Lane<D, LT, J, T>>
struct Walker {
    static __device__ __forceinline__ void run(const Ln& ln, const int* lim, T pp, uint32_t off,
                                               int hsum, int rem, St& st) {
        int b;
        if constexpr (EXPL) {
            const int code = st.bounds[st.bq++];
            if constexpr (J > 0 && K == J - 1) {
                if (code < 0x80) {
                    static_block_dispatch<J, LT, B>(code, ln, pp, off, hsum, st);
                    return;
                }
            }
            b = code & 0x7f;
        } else {
            constexpr int kUp2 = St::kBlk ? blk_up2_max<T>() : up2_max<D, T, St::kStaged>();
            if constexpr (J > 0 && K == J + 1 && K < D && kUp2 > 0) {
                bool fits = rem <= kUp2 && rem <= lim[K] && rem <= lim[K - 1];
#pragma unroll
                for (int j = 0; j < J; ++j) fits = fits && rem <= lim[j];
                if (fits) {
                    up_block_dispatch<K, J, LT, B, kUp2>(rem, ln, pp, off, hsum, st);
                    return;
                }
            }
            constexpr int kUp = up_max<D, T, St::kStaged>();
            if constexpr (J > 0 && K == J && K < D && kUp > 0) {
                bool fits = rem <= kUp && rem <= lim[K];
#pragma unroll
                for (int j = 0; j < J; ++j) fits = fits && rem <= lim[j];
                if (fits) {
                    up_block_dispatch<K, J, LT, B, kUp>(rem, ln, pp, off, hsum, st);
                    return;
                }
            }
            if constexpr (J > 0 && K == J - 1) {
                bool fits = true;
#pragma unroll
                for (int j = 0; j < J; ++j) fits = fits && rem <= lim[j];
                if (fits) {
                    static_block_dispatch<J, LT, B>(rem, ln, pp, off, hsum, st);
                    return;
                }
            }
            b = min(rem, lim[K]);
        }
        // walk dimension K incrementally (outer dims, or a block cut by a restriction)
        T m = ln.m0[K];
        uint32_t t = 0;
        for (int l = 0; l <= b; ++l) {
            const T phi = T(1) - fabs(m);
            const uint32_t o = off + ((t * uint32_t(sizeof(T))) << hsum);
            if constexpr (K == 0) st.leaf(pp * phi, o);
            else Walker<K - 1, D, LT, J, B, EXPL, T, St, Ln>::run(ln, lim, pp * phi, o, hsum + l,
                                                               rem - l, st);
            const bool ge = m >= T(0);
            t = 2u * t + (ge ? 1u : 0u);
            m = T(2) * m + (ge ? T(-1) : T(1));
        }
    }
};

// NM1 >= 0: the block-staged walk of a restricted-regular grid with n - 1 = NM1, in
// workgroups of TPB threads with two stage buffers of STB bytes.
template <int D, int LT, int J, int B, int W, bool EXPL, bool STG, class T, class Acc,
          int NM1 = -1, int TPB = kThreads, int STB = kStageBytes>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1)))
void k_eval(Params<T> prm) {
    static_assert(NM1 < 0 || (STG && !EXPL), "the block-staged walk is a staged regular walk");
    int gi;
    int64_t first, count;
    if (prm.work) {
        const cptr<WorkItem> w = as_const(prm.work) + blockIdx.x;
        gi = w->grid;
        first = w->first;
        count = w->count;
    } else {
        gi = 0;
        first = int64_t(blockIdx.x) * blockDim.x;
        count = prm.P - first < int64_t(blockDim.x) ? prm.P - first : int64_t(blockDim.x);
    }
    const cptr<GridDesc> G = as_const(prm.grids) + gi;
    const int64_t p = first + threadIdx.x;
    const bool live = int64_t(threadIdx.x) < count;
    // idle lanes walk the workgroup's first point and store nothing; with a
    // permutation, neighbouring lanes hold spatially close points, so a wave's gathers
    // share cache lines
    const int64_t pc = live ? (prm.perm ? int64_t(prm.perm[p]) : p) : first;
    constexpr bool kSparse = NM1 >= 0 && J > 0 && D > J && blk_sparse(D, int(sizeof(T)));
    constexpr bool kLean = NM1 >= 0 && !kSparse && blk_lean_walk<T>();
    using Ln = std::conditional_t<kSparse, LaneSp<(J > 0 ? J : 1), T>,
                                  std::conditional_t<kLean, LaneBlk<D, J, T, (W >= 4)>,
                                                     Lane<D, LT, J, T, (NM1 >= 0 && W >= 4)>>>;
    Ln ln;
    // sparse walk: coordinates of dimensions J.. and the level-0 hat products Z_c, c =
    // J..D (single_leaf), as 32-bit words (coord_at)
    constexpr int kXs = kSparse ? D - J : 1;
    uint32_t xlo[kXs], xhi[sizeof(T) == 8 ? kXs : 1];
    uint32_t zlo[kXs + 1], zhi[sizeof(T) == 8 ? kXs + 1 : 1];
    const T* x = prm.coords + G->coord_off + pc * D;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const T xk = x[k];
        if constexpr (J > 0) {
            if (k < J) fill_tables(ln, k < J ? k : 0, xk);
        }
        if constexpr (kSparse) {
            if constexpr (TopOnTheFly<Ln>::value)
                if (k == J - 1) ln.x[J - 1] = xk;
            if (k >= J) {
                if constexpr (sizeof(T) == 8) {
                    const uint64_t b = __builtin_bit_cast(uint64_t, xk);
                    xlo[k >= J ? k - J : 0] = uint32_t(b);
                    xhi[k >= J ? k - J : 0] = uint32_t(b >> 32);
                } else {
                    xlo[k >= J ? k - J : 0] = __builtin_bit_cast(uint32_t, xk);
                }
            }
        } else if constexpr (kLean) {
            ln.x[k] = xk;
        } else {
            ln.m0[k] = T(2) * xk - T(1);
            if constexpr (TopOnTheFly<Ln>::value)
                if (k == J - 1) ln.x[J - 1 >= 0 ? J - 1 : 0] = xk;
        }
    }
    if constexpr (kSparse) {
        T z = T(1);
#pragma unroll
        for (int j = 0; j < J; ++j) z = z * ln.ph[j][0];
#pragma unroll
        for (int i = 0; i <= kXs; ++i) {
            if constexpr (sizeof(T) == 8) {
                const uint64_t b = __builtin_bit_cast(uint64_t, z);
                zlo[i] = uint32_t(b);
                zhi[i] = uint32_t(b >> 32);
            } else {
                zlo[i] = __builtin_bit_cast(uint32_t, z);
            }
            if (i < kXs) {
                T phi0;
                uint32_t t0;
                hat_at(x[J + (i < kXs ? i : 0)], 0, phi0, t0);
                z = z * phi0;
            }
        }
    }
    int lim[D];
    if constexpr (!kLean && !kSparse) {
#pragma unroll
        for (int k = 0; k < D; ++k) lim[k] = G->limits[k];
    }
    // buffer descriptor word 3 for gfx9-family raw (untyped) buffer access
    const T* alpha = prm.alpha ? prm.alpha + G->alpha_off : static_cast<const T*>(G->alpha_ptr);
    using St = std::conditional_t<(NM1 >= 0), StateBlk<T, Acc, NM1, TPB, STB>,
                                  State<T, Acc, STG>>;
    St st{};
    st.alpha = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(alpha), short(0),
                                                 int(G->N * int64_t(sizeof(T))), 0x00020000);
    st.leaf_base = as_const(G->leaf_base);
    st.bounds = as_const(G->bounds);
    st.q = 0;
    st.bq = 0;
    if constexpr (NM1 >= 0) {
        __shared__ __attribute__((aligned(16))) uint8_t blk_buf[2 * STB];
        const cptr<GridDesc::BlkStages> tb = &G->blk[sizeof(T) == 8 ? 0 : 1];
        st.lds = (lds_byte*)blk_buf;
        st.seg_begin = as_const(tb->seg_begin);
        st.segs = as_const(tb->segs);
        st.leaf_base = as_const(tb->leaf_base);
        st.num_stages = tb->num_stages;
        st.stage = 0;
        st.base = 0;
        st.cur = 0;
        st.limit = uint32_t(STB);
        st.begin();
    } else if constexpr (STG) {
        __shared__ __attribute__((aligned(16))) uint8_t stage_buf[2 * kStageBytes];
        const cptr<GridDesc::Stages> tb = &G->stage[sizeof(T) == 8 ? 0 : 1];
        st.lds = (lds_byte*)stage_buf;
        st.leaf_lds = as_const(tb->leaf_lds);
        st.stage_first = as_const(tb->stage_first);
        st.seg_begin = as_const(tb->seg_begin);
        st.segs = as_const(tb->segs);
        st.num_stages = tb->num_stages;
        st.stage = 0;
        st.next_first = -1;
        if (st.num_stages > 0) {
            st.issue(0);
            stage_barrier();
            st.next_first = st.stage_first[1];
            if (st.num_stages > 1) st.issue(1);
        }
    }
    if constexpr (kSparse)
        WalkerSp<0, (NM1 > 0 ? NM1 : 0), J, B, T, St, Ln>::run(ln, xlo, xhi, zlo, zhi, D, T(1),
                                                               0u, 0, NM1, st);
    else if constexpr (kLean)
        WalkerBlk<D - 1, D, J, B, T, St, Ln>::run(ln, T(1), 0u, 0, NM1, st);
    else
        Walker<D - 1, D, LT, J, B, EXPL, T, St, Ln>::run(ln, lim, T(1), 0u, 0, G->n - 1, st);
    if (live) prm.out[G->out_off + pc] = st.acc.value();
}

// The block-staged lean walk with two points per lane (variant 5; one-grid launches of a
// regular grid with n - 1 = NM1): workgroup w evaluates points w 2 TPB + i TPB + lane,
// i = 0, 1 (coalesced coordinate loads per half).  Idle points walk the workgroup's first
// point and store nothing.
template <int D, int J, int B, int W, class T, class Acc, int NM1, int TPB, int STB>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(W)))
void k_eval2(Params<T> prm) {
    static_assert(NM1 >= 0 && J > 0 && J < D, "a regular block-staged walk");
    const cptr<GridDesc> G = as_const(prm.grids);
    const int64_t first = int64_t(blockIdx.x) * (2 * TPB);
    LaneBlk2<D, J, T> ln;
    int64_t pc[2];
    bool live[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int64_t p = first + i * TPB + threadIdx.x;
        live[i] = p < prm.P;
        pc[i] = live[i] ? (prm.perm ? int64_t(prm.perm[p]) : p) : first;
        const T* x = prm.coords + G->coord_off + pc[i] * D;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const T xk = x[k];
            if (k < J) fill_tables(ln.p[i], k < J ? k : 0, xk);
            ln.p[i].x[k] = xk;
        }
    }
    const T* alpha = prm.alpha ? prm.alpha + G->alpha_off : static_cast<const T*>(G->alpha_ptr);
    using St = StateBlk<T, AccPair<Acc>, NM1, TPB, STB>;
    St st{};
    st.alpha = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(alpha), short(0),
                                                 int(G->N * int64_t(sizeof(T))), 0x00020000);
    st.bounds = as_const(G->bounds);
    st.q = 0;
    st.bq = 0;
    __shared__ __attribute__((aligned(16))) uint8_t blk_buf[2 * STB];
    const cptr<GridDesc::BlkStages> tb = &G->blk[sizeof(T) == 8 ? 0 : 1];
    st.lds = (lds_byte*)blk_buf;
    st.seg_begin = as_const(tb->seg_begin);
    st.segs = as_const(tb->segs);
    st.leaf_base = as_const(tb->leaf_base);
    st.num_stages = tb->num_stages;
    st.stage = 0;
    st.base = 0;
    st.cur = 0;
    st.limit = uint32_t(STB);
    st.begin();
    WalkerBlk2<D - 1, D, J, B, T, St>::run(ln, T(1), T(1), 0u, 0u, 0, NM1, st);
#pragma unroll
    for (int i = 0; i < 2; ++i)
        if (live[i]) prm.out[G->out_off + pc[i]] = st.acc.a[i].value();
}

// The staged-pattern probe: the data path of a staged walk without its arithmetic.  The
// workgroup runs the walk's stage plan (the same buffer-to-LDS copies and stage barriers
// between them); in stage k every lane reads as many coefficients from LDS as the walk
// reads there (the stage's staged leaves), at random cells of the stage, and adds each
// with one FMA; then it reads the leaves the walk gathers from global memory, one random
// cell of each leaf's block.  No hats, no walk, no per-leaf addressing: it does strictly
// less work per read than the walk, so its time is a floor for the walk's on the grid.
template <class T>
struct ProbeParams {
    const T* alpha;
    int64_t N;
    const int32_t* info;         // [K][2]: staged leaves of stage k, cell mask of the stage
    const int32_t* glob;         // [G][2]: base offset and level sum of each gathered leaf
    const int32_t* seg_begin;
    const int32_t* segs;
    int num_stages;
    int num_glob;
    int64_t P;
    uint32_t seed;
    T* out;
};

// PPL points per lane (the two-point walk's data path: each stage serves 2 TPB points)
template <class T, int TPB, int STB, int PPL = 1>
__global__ __launch_bounds__(TPB) void k_stage_probe(ProbeParams<T> p) {
    __shared__ __attribute__((aligned(16))) uint8_t buf[2 * STB];
    lds_byte* lds = (lds_byte*)buf;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T*>(p.alpha), short(0), int(p.N * int64_t(sizeof(T))), 0x00020000);
    const cptr<int32_t> info = as_const(p.info), glob = as_const(p.glob),
                        seg_begin = as_const(p.seg_begin), segs = as_const(p.segs);
    int64_t pi[PPL];
    uint32_t h[PPL];
    T w[PPL], acc[PPL];
#pragma unroll
    for (int r = 0; r < PPL; ++r) {
        pi[r] = int64_t(blockIdx.x) * (PPL * TPB) + r * TPB + threadIdx.x;
        uint32_t hr = uint32_t(pi[r]) * 0x9E3779B9u ^ p.seed;
        hr ^= hr >> 15;
        hr *= 0x85EBCA6Bu;
        hr ^= hr >> 13;
        h[r] = hr;
        w[r] = T(1) + T(hr & 1023u) * T(1.0 / 4096);
        acc[r] = T(0);
    }
    const int K = p.num_stages;
    if (K > 0) {
        copy_stage<TPB>(rsrc, lds, seg_begin, segs, 0);
        stage_barrier();
        if (K > 1) copy_stage<TPB>(rsrc, lds, seg_begin, segs, 1);
    }
    for (int k = 0; k < K; ++k) {
        if (k > 0) {
            stage_barrier();
            if (k + 1 < K) copy_stage<TPB>(rsrc, lds, seg_begin, segs, k + 1);
        }
        const int cnt = info[2 * k];
        const uint32_t mask = uint32_t(info[2 * k + 1]);
        const uint32_t base = uint32_t(k & 1) * uint32_t(STB);
#pragma unroll
        for (int r = 0; r < PPL; ++r) {
            uint32_t at[8];
#pragma unroll
            for (int j = 0; j < 8; ++j)
                at[j] = base + (((h[r] >> j) ^ (h[r] >> (j + 11))) & mask) * uint32_t(sizeof(T));
            int i = 0;
            for (; i + 8 <= cnt; i += 8) {
                asm volatile("" ::: "memory");     // re-read every pass (no load hoisting)
                T v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    v[j] = *reinterpret_cast<const __attribute__((address_space(3))) T*>(lds + at[j]);
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[r] = fma(v[j], w[r], acc[r]);
            }
            for (; i < cnt; ++i) {
                asm volatile("" ::: "memory");
                acc[r] = fma(*reinterpret_cast<const __attribute__((address_space(3))) T*>(
                                 lds + at[i & 7]),
                             w[r], acc[r]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < PPL; ++r) {
        for (int q = 0; q < p.num_glob; ++q) {
            const uint32_t cell = (h[r] >> (q & 7)) & ((1u << glob[2 * q + 1]) - 1u);
            acc[r] = fma(gather_leaf<T>(rsrc, glob[2 * q], cell * uint32_t(sizeof(T))), w[r], acc[r]);
        }
        if (pi[r] < p.P) p.out[pi[r]] = acc[r];
    }
}

// Launch the staged-pattern probe of grid g's staged walk in this precision (the
// block-staged one where the grid runs it, else the staged one).
template <class T>
int launch_stage_probe(const Grid& g, const DeviceTables& t, int variant, const T* alpha,
                       int64_t P, uint32_t seed, T* out, hipStream_t s);

// Compiled variants of the register-table walk (LT = 8); the autotuner picks one per
// grid shape and precision.  Tables for J dimensions cost 3 J LT registers (FP64);
// larger batches keep more gathers in flight per wave, fewer registers more waves.
struct Variant {
    int J, B, W;
    bool staged;
};
constexpr Variant kVariants[] = {{3, 8, 3, false}, {2, 8, 3, false}, {3, 4, 4, false},
                                 {kStageJ, 0, 3, true},     // batch: stage_batch()
                                 {kStageJ, 0, 3, true},     // waves: blk_waves(d, n - 1)
                                 {kStageJ, 0, 2, true}};    // two points per lane
static_assert(kStagedVariant == 3 && kVariants[kStagedVariant].staged, "staged variant index");
static_assert(kBlkVariant == 4 && kVariants[kBlkVariant].staged, "block-staged variant index");
static_assert(kBlk2Variant == 5 && kVariants[kBlk2Variant].staged, "two-point variant index");
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
// tabled dimensions the explicit-set kernels use (their node-code stream depends on it)
constexpr int kExplicitJ = kExplicitTabDims;

template <int D, int LT, int J, int B, int W, bool EXPL, class T, class Acc, bool STG = false>
hipError_t launch_one(const Params<T>& prm, int64_t blocks, int block, hipStream_t s) {
    constexpr int Jd = J < D ? J : D;
    hipLaunchKernelGGL((k_eval<D, LT, Jd, B, W, EXPL, STG, T, Acc>), dim3(unsigned(blocks)),
                       dim3(block), 0, s, prm);
    return hipGetLastError();
}

// variant index -> (J, B, W) with LT = 8, or the table-free walk with LT = 16.
// Instantiated per dimension range in sg_eval_inst.hip (one object per range and
// precision, compiled in parallel).
template <int D, bool EXPL, class T, class Acc>
hipError_t launch_variant(int variant, int lt, const Params<T>& prm, int64_t blocks, int block,
                          hipStream_t s)
#ifdef SG_EVAL_INSTANTIATE
{
    if (lt > 8) return launch_one<D, 16, 0, 8, 0, EXPL, T, Acc>(prm, blocks, block, s);
    if constexpr (EXPL) {
        static_assert(kExplicitJ == kStageJ, "explicit walk and stage plan table alike");
        if (variant == kStagedVariant)
            return launch_one<D, 8, kExplicitJ, stage_batch<D, T>(), 3, true, T, Acc, true>(prm, blocks,
                                                                                 block, s);
        return launch_one<D, 8, kExplicitJ, 8, 3, true, T, Acc>(prm, blocks, block, s);
    } else {
        switch (variant) {
            case 1: return launch_one<D, 8, 2, 8, 3, false, T, Acc>(prm, blocks, block, s);
            case 2: return launch_one<D, 8, 3, 4, 4, false, T, Acc>(prm, blocks, block, s);
            case kStagedVariant:
                return launch_one<D, 8, kStageJ, stage_batch<D, T>(), 3, false, T, Acc, true>(
                    prm, blocks, block, s);
            default: return launch_one<D, 8, 3, 8, 3, false, T, Acc>(prm, blocks, block, s);
        }
    }
}
#else
;
#endif

// The block-staged lean walk with two points per lane (FP64, depth-first order), 256
// threads and 512 points per workgroup, 2 waves per SIMD (its lane state for two points
// needs the 256 VGPRs of that occupancy).
constexpr int kBlk2Threads = 256;
#ifndef SG_BLK2_BATCH
#define SG_BLK2_BATCH 2
#endif
template <int D, class T, class Acc>
hipError_t launch_blk2(int nm1, const Params<T>& prm, int64_t blocks, hipStream_t s)
#ifdef SG_EVAL_INSTANTIATE
{
    constexpr int Jd = kStageJ < D ? kStageJ : D;
    constexpr int E = int(sizeof(T));
    if constexpr (blk2_built(D, E) && Jd < D) {
        switch (nm1) {
#define SG_NM2(NN)                                                                           \
    case NN:                                                                                 \
        hipLaunchKernelGGL((k_eval2<D, Jd, SG_BLK2_BATCH, 2, T, Acc, NN, kBlk2Threads,        \
                                    blk_stage_bytes(D, NN, E)>),                             \
                           dim3(unsigned(blocks)), dim3(kBlk2Threads), 0, s, prm);           \
        return hipGetLastError();
            SG_NM2(2) SG_NM2(3) SG_NM2(4) SG_NM2(5) SG_NM2(6) SG_NM2(7)
#undef SG_NM2
            default: break;
        }
    }
    return hipErrorInvalidValue;
}
#else
;
#endif

template <class T, class Acc>
hipError_t launch_d_blk2(int d, int nm1, const Params<T>& prm, int64_t blocks, hipStream_t s) {
    switch (d) {
#define SG_D(DD) \
    case DD: return launch_blk2<DD, T, Acc>(nm1, prm, blocks, s);
        SG_D(1) SG_D(2) SG_D(3) SG_D(4) SG_D(5) SG_D(6) SG_D(7) SG_D(8) SG_D(9) SG_D(10)
        SG_D(11) SG_D(12) SG_D(13) SG_D(14) SG_D(15) SG_D(16) SG_D(17) SG_D(18) SG_D(19)
        SG_D(20)
#undef SG_D
        default: return hipErrorInvalidValue;
    }
}

// The block-staged walk of a regular grid with n - 1 = nm1 (kBlkMinNM1 .. kBlkMaxNM1),
// workgroups of blk_threads(D, nm1, sizeof(T)); built for the dimensions blk_built names.
template <int D, class T, class Acc>
hipError_t launch_blk(int nm1, const Params<T>& prm, int64_t blocks, hipStream_t s)
#ifdef SG_EVAL_INSTANTIATE
{
    constexpr int Jd = kStageJ < D ? kStageJ : D;
    constexpr int E = int(sizeof(T));
    if constexpr (blk_built(D, E)) {
        switch (nm1) {
#define SG_NM(NN)                                                                          \
    case NN:                                                                               \
        hipLaunchKernelGGL((k_eval<D, 8, Jd, blk_batch<D, T>(), blk_waves(D, NN, E), false,  \
                                   true, T, Acc, NN, blk_threads(D, NN, E),                 \
                                   blk_stage_bytes(D, NN, E)>),                             \
                           dim3(unsigned(blocks)), dim3(blk_threads(D, NN, E)), 0, s, prm); \
        return hipGetLastError();
            SG_NM(2) SG_NM(3) SG_NM(4) SG_NM(5) SG_NM(6) SG_NM(7)
#undef SG_NM
            default: break;
        }
    }
    return hipErrorInvalidValue;
}
#else
;
#endif

template <class T, class Acc>
hipError_t launch_d_blk(int d, int nm1, const Params<T>& prm, int64_t blocks, hipStream_t s) {
    switch (d) {
#define SG_D(DD) \
    case DD: return launch_blk<DD, T, Acc>(nm1, prm, blocks, s);
        SG_D(1) SG_D(2) SG_D(3) SG_D(4) SG_D(5) SG_D(6) SG_D(7) SG_D(8) SG_D(9) SG_D(10)
        SG_D(11) SG_D(12) SG_D(13) SG_D(14) SG_D(15) SG_D(16) SG_D(17) SG_D(18) SG_D(19)
        SG_D(20)
#undef SG_D
        default: return hipErrorInvalidValue;
    }
}

template <bool EXPL, class T, class Acc>
hipError_t launch_d(int d, int variant, int lt, const Params<T>& prm, int64_t blocks, int block,
                    hipStream_t s) {
    switch (d) {
#define SG_D(DD) \
    case DD: return launch_variant<DD, EXPL, T, Acc>(variant, lt, prm, blocks, block, s);
        SG_D(1) SG_D(2) SG_D(3) SG_D(4) SG_D(5) SG_D(6) SG_D(7) SG_D(8) SG_D(9) SG_D(10)
        SG_D(11) SG_D(12) SG_D(13) SG_D(14) SG_D(15) SG_D(16) SG_D(17) SG_D(18) SG_D(19)
        SG_D(20)
#undef SG_D
        default: return hipErrorInvalidValue;
    }
}

}  // namespace eval
}  // namespace sg

namespace sg {
namespace eval {

// Walk variant and table length for grid g: the J tabled dims need levels below 8.
inline int table_levels(const Grid& g, int variant) {
    if (g.explicit_set) return explicit_tab_dims(g) == 0 ? 16 : 8;
    const int J = kVariants[variant].J < g.d ? kVariants[variant].J : g.d;
    for (int j = 0; j < J; ++j)
        if (g.limits[j] > 7) return 16;
    return 8;
}

template <class T, class Acc>
int evaluate_launch(const Grid& g, const DeviceTables& t, const T* alpha, const T* coords,
                    int64_t P, T* out, const EvalTuning& tu, const uint32_t* perm,
                    hipStream_t s) {
    if (P == 0) return kOk;
    if (tu.block != 64 && tu.block != 128 && tu.block != 256) {
        set_error("evaluate: block must be 64, 128 or 256");
        return kInvalid;
    }
    if (tu.variant < 0 || tu.variant >= kNumVariants) {
        set_error("evaluate: unknown kernel variant");
        return kInvalid;
    }
    if (g.N * int64_t(sizeof(T)) > int64_t(0x7fffffff)) {
        set_error("evaluate: surpluses beyond 2 GiB");
        return kRange;
    }
    const int v = effective_variant(g, tu.variant, int(sizeof(T)));
    const int block = v == kBlkVariant ? blk_threads(g.d, g.n - 1, int(sizeof(T)))
                      : kVariants[v].staged ? kThreads : tu.block;
    const int points_per_block = v == kBlk2Variant ? 2 * kBlk2Threads : block;
    const int64_t blocks = (P + points_per_block - 1) / points_per_block;
    if (blocks > int64_t(0x7fffffff)) {
        set_error("evaluate: too many points for one launch");
        return kRange;
    }
    Params<T> prm{};
    prm.alpha = alpha;
    prm.coords = coords;
    prm.out = out;
    prm.grids = static_cast<const GridDesc*>(t.desc);
    prm.work = nullptr;
    prm.perm = perm;
    prm.P = P;
    const hipError_t e = v == kBlk2Variant
        ? launch_d_blk2<T, Acc>(g.d, g.n - 1, prm, blocks, s)
        : v == kBlkVariant
        ? launch_d_blk<T, Acc>(g.d, g.n - 1, prm, blocks, s)
        : g.explicit_set
        ? launch_d<true, T, Acc>(g.d, v == kStagedVariant ? v : 0, table_levels(g, 0), prm,
                                 blocks, block, s)
        : launch_d<false, T, Acc>(g.d, v, table_levels(g, v), prm, blocks, block, s);
    if (e != hipSuccess) {
        set_error(std::string("k_eval launch: ") + hipGetErrorString(e));
        return kHip;
    }
    return kOk;
}

template <class T>
int launch_stage_probe(const Grid& g, const DeviceTables& t, int variant, const T* alpha,
                       int64_t P, uint32_t seed, T* out, hipStream_t s) {
    if (P == 0) return kOk;
    const int e = sizeof(T) == 8 ? 0 : 1;
    const int v = effective_variant(g, variant, int(sizeof(T)));
    if (v != kBlkVariant && v != kBlk2Variant && v != kStagedVariant) {
        set_error("stage probe: the grid runs no staged walk in this precision");
        return kUnsupported;
    }
    const GridDesc& hd = t.host_desc;
    const DeviceTables::Probe& pt = t.probe[v == kStagedVariant ? 0 : 1][e];
    ProbeParams<T> p{};
    p.alpha = alpha;
    p.N = g.N;
    p.info = pt.info;
    p.glob = pt.glob;
    p.num_stages = pt.num_stages;
    p.num_glob = pt.num_glob;
    p.P = P;
    p.seed = seed;
    p.out = out;
    // the block-staged plans' stage size, and the walk's workgroup, follow the shape
    const bool quad = blk_quad(g.d, g.n - 1, int(sizeof(T)));
    if (v == kBlk2Variant) {
        p.seg_begin = hd.blk[e].seg_begin;
        p.segs = hd.blk[e].segs;
        constexpr int TPB = kBlk2Threads;
        const dim3 grid(unsigned((P + 2 * TPB - 1) / (2 * TPB)));
        if (quad)
            hipLaunchKernelGGL((k_stage_probe<T, TPB, kBlkQuadStageBytes, 2>), grid, dim3(TPB),
                               0, s, p);
        else
            hipLaunchKernelGGL((k_stage_probe<T, TPB, kBlkStageBytes, 2>), grid, dim3(TPB), 0, s,
                               p);
    } else if (v == kBlkVariant) {
        p.seg_begin = hd.blk[e].seg_begin;
        p.segs = hd.blk[e].segs;
        if (quad) {
            constexpr int TPB = kBlkQuadThreads;
            hipLaunchKernelGGL((k_stage_probe<T, TPB, kBlkQuadStageBytes>),
                               dim3(unsigned((P + TPB - 1) / TPB)), dim3(TPB), 0, s, p);
        } else {
            constexpr int TPB = kBlkThreads;
            hipLaunchKernelGGL((k_stage_probe<T, TPB, kBlkStageBytes>),
                               dim3(unsigned((P + TPB - 1) / TPB)), dim3(TPB), 0, s, p);
        }
    } else {
        p.seg_begin = hd.stage[e].seg_begin;
        p.segs = hd.stage[e].segs;
        hipLaunchKernelGGL((k_stage_probe<T, kThreads, kStageBytes>),
                           dim3(unsigned((P + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, p);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_error(std::string("stage probe launch: ") + hipGetErrorString(err));
        return kHip;
    }
    return kOk;
}

}  // namespace eval
}  // namespace sg
