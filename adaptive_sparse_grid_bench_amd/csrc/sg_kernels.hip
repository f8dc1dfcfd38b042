// HIP kernels of the compact sparse grid hot path, written for gfx950 (CDNA4, wave64).
//
//   k_idx2gp / k_gp2idx   the offset <-> (levels, indices) bijection, asg.c:187-243
//   k_grid_coords         grid-point coordinates (2i+1)/2^(l+1) per offset, asg.c:143-152
//   k_hierarchize_dim     one dimension's sweep v -= (left + right)/2 over pole families
//                         held in LDS, asg.c:557-571 semantics (k_hierarchize_dim_global:
//                         the same in global memory for families too large for LDS)
//   k_eval_dfs            interpolant at one point per lane, depth-first walk over the
//                         level set with incremental hat products, asg.c:1237-1274 values
//   k_eval_ref            the same interpolant in the reference's exact operation order
//                         (bit-identical to evaluate0)
//   k_uniform             seeded query points k * 2^-24 (splitmix64 counter stream)
//
// ("asg.c" = adaptive_sparse_grid_bench.c of the reference benchmark.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "sg_kernels.hpp"

namespace sg {

namespace {

constexpr int kThreads = 256;
// LDS of a pole-sweep tile (values and family starts); larger tiles sweep in global memory
constexpr size_t kPoleLdsBytes = 64 * 1024;

#define SG_HIP_CHECK(expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));         \
            return kHip;                                                          \
        }                                                                         \
    } while (0)

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return kHip;
    }
    return kOk;
}

// --------------------------------------------------------------------- bijection --

struct BijParams {
    int d;
    int n;
    int explicit_set;
    int64_t N;
    const int32_t* dp;          // [d][n]
    const int32_t* lidx;        // [n+1]
    const int32_t* cnt;         // [n+1]
    const uint8_t* sub_levels;  // [S][d]
    const int32_t* sub_base;    // [S]
    int limits[kMaxDims];
};

constexpr int kMaxN = kMaxLevelSum + 1;

__device__ __forceinline__ void decode_offset(const BijParams& p, const int32_t* sdp,
                                              const int32_t* slidx, const int32_t* scnt,
                                              int64_t off, int32_t* lv, int32_t* iv) {
    int s = 0;
    while (s + 1 < p.n && off >= slidx[s + 1]) ++s;
    int64_t r = off - slidx[s];
    int64_t rank = r >> s;
    uint32_t bits = static_cast<uint32_t>(r & ((int64_t(1) << s) - 1));
    if (p.explicit_set) {
        const uint8_t* l = p.sub_levels + (int64_t(scnt[s]) + rank) * p.d;
        for (int i = 0; i < p.d; ++i) lv[i] = l[i];
    } else {
        // greedy descent through dp, asg.c:221-236
        int sum = s;
        for (int i = p.d - 1; i >= 1; --i) {
            int l = 0;
            while (l <= sum && rank >= sdp[(i - 1) * p.n + (sum - l)]) {
                rank -= sdp[(i - 1) * p.n + (sum - l)];
                ++l;
            }
            lv[i] = l;
            sum -= l;
        }
        lv[0] = sum;
    }
    for (int i = p.d - 1; i >= 0; --i) {
        iv[i] = static_cast<int32_t>(bits & ((1u << lv[i]) - 1u));
        bits >>= lv[i];
    }
}

__device__ __forceinline__ void load_bij_tables(const BijParams& p, int32_t* sdp,
                                                int32_t* slidx, int32_t* scnt) {
    if (!p.explicit_set)
        for (int i = threadIdx.x; i < p.d * p.n; i += blockDim.x) sdp[i] = p.dp[i];
    for (int i = threadIdx.x; i <= p.n; i += blockDim.x) {
        slidx[i] = p.lidx[i];
        scnt[i] = p.cnt[i];
    }
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_idx2gp(BijParams p, const int32_t* index,
                                                     int64_t count, int32_t* levels,
                                                     int32_t* indices) {
    __shared__ int32_t sdp[kMaxDims * kMaxN];
    __shared__ int32_t slidx[kMaxN + 1];
    __shared__ int32_t scnt[kMaxN + 1];
    load_bij_tables(p, sdp, slidx, scnt);
    int32_t lv[kMaxDims], iv[kMaxDims];
    for (int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < count;
         o += int64_t(gridDim.x) * blockDim.x) {
        int64_t off = index ? int64_t(index[o]) : o;
        int32_t* lo = levels + o * p.d;
        int32_t* io = indices + o * p.d;
        if (off < 0 || off >= p.N) {
            for (int i = 0; i < p.d; ++i) { lo[i] = -1; io[i] = -1; }
            continue;
        }
        decode_offset(p, sdp, slidx, scnt, off, lv, iv);
        for (int i = 0; i < p.d; ++i) { lo[i] = lv[i]; io[i] = iv[i]; }
    }
}

// compact-order comparison of a level vector against subspace q (explicit sets)
__device__ __forceinline__ int cmp_levels(const int32_t* lv, const uint8_t* l, int d) {
    for (int j = d - 1; j >= 0; --j) {
        if (lv[j] != l[j]) return lv[j] < l[j] ? -1 : 1;
    }
    return 0;
}

__global__ __launch_bounds__(kThreads) void k_gp2idx(BijParams p, const int32_t* levels,
                                                     const int32_t* indices, int64_t count,
                                                     int32_t* index) {
    __shared__ int32_t sdp[kMaxDims * kMaxN];
    __shared__ int32_t slidx[kMaxN + 1];
    __shared__ int32_t scnt[kMaxN + 1];
    load_bij_tables(p, sdp, slidx, scnt);
    int32_t lv[kMaxDims];
    for (int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < count;
         o += int64_t(gridDim.x) * blockDim.x) {
        const int32_t* li = levels + o * p.d;
        const int32_t* ii = indices + o * p.d;
        int sum = 0;
        bool ok = true;
        uint32_t in_block = 0;
        for (int j = 0; j < p.d; ++j) {
            int l = li[j], i = ii[j];
            lv[j] = l;
            ok = ok && l >= 0 && l <= p.limits[j] && i >= 0 && i < (1 << l);
            sum += l;
            in_block = (in_block << (ok ? l : 0)) + uint32_t(ok ? i : 0);
        }
        ok = ok && sum < p.n;
        if (!ok) { index[o] = -1; continue; }
        int64_t rank;
        if (p.explicit_set) {
            int64_t lo = scnt[sum], hi = scnt[sum + 1];   // binary search in [lo, hi)
            while (lo < hi) {
                int64_t mid = (lo + hi) >> 1;
                if (cmp_levels(lv, p.sub_levels + mid * p.d, p.d) > 0) lo = mid + 1;
                else hi = mid;
            }
            if (lo >= scnt[sum + 1] || cmp_levels(lv, p.sub_levels + lo * p.d, p.d) != 0) {
                index[o] = -1;
                continue;
            }
            rank = lo - scnt[sum];
        } else {
            // asg.c:195-201
            rank = 0;
            int part = lv[0];
            for (int i = 1; i < p.d; ++i) {
                part += lv[i];
                for (int j = 0; j < lv[i]; ++j) rank += sdp[(i - 1) * p.n + (part - j)];
            }
        }
        index[o] = static_cast<int32_t>(slidx[sum] + (rank << sum) + in_block);
    }
}

__global__ __launch_bounds__(kThreads) void k_grid_coords(BijParams p, double* coords) {
    __shared__ int32_t sdp[kMaxDims * kMaxN];
    __shared__ int32_t slidx[kMaxN + 1];
    __shared__ int32_t scnt[kMaxN + 1];
    load_bij_tables(p, sdp, slidx, scnt);
    int32_t lv[kMaxDims], iv[kMaxDims];
    for (int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < p.N;
         o += int64_t(gridDim.x) * blockDim.x) {
        decode_offset(p, sdp, slidx, scnt, o, lv, iv);
        for (int i = 0; i < p.d; ++i)
            coords[o * p.d + i] = ldexp(2.0 * iv[i] + 1.0, -(lv[i] + 1));
    }
}

// asg.c:87-94 and its coordinates asg.c:148-151, in T, left to right, no contraction:
// with T = float the value is bit-identical to the reference's init().
template <class T>
__device__ __forceinline__ T fct_reference(int d, const int32_t* lv, const int32_t* iv) {
#pragma clang fp contract(off)
    T res = T(0);
    for (int k = 0; k < d; ++k) {
        const T x = T(iv[k]) * T(1) / T(1u << lv[k]) + T(1) / T(1u << (lv[k] + 1));
        res += x * x * (T(1) - x) * (T(1) - x);
    }
    return res;
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_sample(BijParams p, SampleParams sp, T* out) {
    __shared__ int32_t sdp[kMaxDims * kMaxN];
    __shared__ int32_t slidx[kMaxN + 1];
    __shared__ int32_t scnt[kMaxN + 1];
    load_bij_tables(p, sdp, slidx, scnt);
    int32_t lv[kMaxDims], iv[kMaxDims];
    for (int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < p.N;
         o += int64_t(gridDim.x) * blockDim.x) {
        decode_offset(p, sdp, slidx, scnt, o, lv, iv);
        T v;
        if (sp.fn == kFnReference) {
            v = fct_reference<T>(p.d, lv, iv);
        } else {
            // coordinates (2 i + 1) / 2^(l+1) are exact in T
            double acc = sp.fn == kFnGaussianSum ? 0.0 : 1.0;
            if (sp.fn == kFnGauss) {
                double e = 0.0;
                for (int k = 0; k < p.d; ++k) {
                    const double x = ldexp(2.0 * iv[k] + 1.0, -(lv[k] + 1)) - sp.centre[0][k];
                    e += x * x;
                }
                acc = exp(-e);
            } else if (sp.fn == kFnGaussianSum) {
                for (int g = 0; g < sp.terms; ++g) {
                    double e = 0.0;
                    for (int k = 0; k < p.d; ++k) {
                        const double x = ldexp(2.0 * iv[k] + 1.0, -(lv[k] + 1)) - sp.centre[g][k];
                        e += x * x / (2.0 * sp.width[g][k] * sp.width[g][k]);
                    }
                    acc += exp(-e);
                }
            } else {
                for (int k = 0; k < p.d; ++k) {
                    const double x = ldexp(2.0 * iv[k] + 1.0, -(lv[k] + 1));
                    acc *= sp.fn == kFnParabola ? 4.0 * x * (1.0 - x) : sin(M_PI * x);
                }
                if (sp.fn == kFnSinProduct) acc *= sp.scale;
            }
            v = T(acc);
        }
        out[o] = v;
    }
}

BijParams make_bij(const Grid& g, const DeviceTables& t) {
    BijParams p{};
    p.d = g.d;
    p.n = g.n;
    p.explicit_set = g.explicit_set ? 1 : 0;
    p.N = g.N;
    p.dp = t.dp;
    p.lidx = t.lidx;
    p.cnt = t.cnt;
    p.sub_levels = t.sub_levels;
    p.sub_base = t.sub_base;
    for (int j = 0; j < g.d; ++j) p.limits[j] = g.limits[j];
    return p;
}

int grid_for(int64_t work) {
    int64_t b = (work + kThreads - 1) / kThreads;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(b, 65536)));
}

// ----------------------------------------------------------------- hierarchize --

// One workgroup per tile of pole families of dimension k.  A family is the set of
// subspaces that share every level but l_k; its blocks (l_k = 0..m) are loaded into
// LDS one after another, so block j of a family with c = sum of the other levels
// starts at 2^c (2^j - 1) and holds 2^(c+j) values.  Inside block j the in-block index
// splits as hi | i_k | lo with widths (c - B, j, B), B = sum of the levels of dims > k
// (asg.c:893-913).  Every point's parents in dimension k (asg.c:828-854) lie in the same
// family, so the sweep reads only LDS and writes each value once, in place.
template <class T>
__global__ __launch_bounds__(kThreads) void k_hierarchize_dim(T* __restrict__ v,
                                                              const int32_t* __restrict__ tiles,
                                                              const int32_t* __restrict__ fams,
                                                              const int32_t* __restrict__ blocks,
                                                              int lds_values) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* buf = reinterpret_cast<T*>(smem);
    int32_t* fstart = reinterpret_cast<int32_t*>(smem + size_t(lds_values) * sizeof(T));
    const int f0 = tiles[3 * blockIdx.x];
    const int nf = tiles[3 * blockIdx.x + 1];
    const int nv = tiles[3 * blockIdx.x + 2];
    for (int i = threadIdx.x; i < nf; i += blockDim.x) fstart[i] = fams[3 * (f0 + i)];
    __syncthreads();

    auto locate = [&](int val, int& f, int& j, int& r, int& c, int& B, int& bfirst) {
        int lo = 0, hi = nf - 1;           // last family whose start is <= val
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (fstart[mid] <= val) lo = mid;
            else hi = mid - 1;
        }
        f = lo;
        const int meta = fams[3 * (f0 + f) + 1];
        bfirst = fams[3 * (f0 + f) + 2];
        c = (meta >> 8) & 255;
        B = meta >> 16;
        const int local = val - fstart[f];
        j = 31 - __clz((local >> c) + 1);
        r = local - (((1 << j) - 1) << c);
    };

    for (int val = threadIdx.x; val < nv; val += blockDim.x) {
        int f, j, r, c, B, bfirst;
        locate(val, f, j, r, c, B, bfirst);
        buf[val] = v[blocks[bfirst + j] + r];
    }
    __syncthreads();
    for (int val = threadIdx.x; val < nv; val += blockDim.x) {
        int f, j, r, c, B, bfirst;
        locate(val, f, j, r, c, B, bfirst);
        if (j == 0) continue;                  // l_k = 0: no parents in dimension k
        const int fbase = fstart[f];
        const int lo = r & ((1 << B) - 1);
        const int ik = (r >> B) & ((1 << j) - 1);
        const int hi = r >> (B + j);
        T left = T(0), right = T(0);
        if (ik != 0) {                          // left parent, asg.c:828-840
            const int tz = __builtin_ctz(ik);
            const int pl = j - 1 - tz;
            const int pi = ik >> (tz + 1);
            left = buf[fbase + (((1 << pl) - 1) << c) + (hi << (B + pl)) + (pi << B) + lo];
        }
        if (ik != (1 << j) - 1) {               // right parent, asg.c:842-854
            const int tz = __builtin_ctz(ik + 1);
            const int pl = j - 1 - tz;
            const int pi = (ik + 1) >> (tz + 1);
            right = buf[fbase + (((1 << pl) - 1) << c) + (hi << (B + pl)) + (pi << B) + lo];
        }
        v[blocks[bfirst + j] + r] = buf[val] - (left + right) * T(0.5);
    }
}

// The same sweep for tiles too large for LDS (over kPoleLdsBytes: a family of more than
// about 8,000 FP64 values, e.g. one dimension reaching level 13 with every other level
// 0), in place in global memory: one workgroup per tile, its families one after another, each family's blocks from the
// finest (j = m) to the coarsest (j = 1) with a workgroup barrier between them.  A
// value's parents in dimension k lie in coarser blocks, which still hold their nodal
// values when it is updated, so every value gets the LDS sweep's expression on the same
// operands, bit for bit.
template <class T>
__global__ __launch_bounds__(kThreads) void k_hierarchize_dim_global(
    T* __restrict__ v, const int32_t* __restrict__ tiles, const int32_t* __restrict__ fams,
    const int32_t* __restrict__ blocks) {
    const int f0 = tiles[3 * blockIdx.x];
    const int nf = tiles[3 * blockIdx.x + 1];
    for (int f = 0; f < nf; ++f) {
        const int meta = fams[3 * (f0 + f) + 1];
        const int bfirst = fams[3 * (f0 + f) + 2];
        const int m = meta & 255, c = (meta >> 8) & 255, B = meta >> 16;
        for (int j = m; j >= 1; --j) {
            const int cnt = 1 << (c + j);          // values of block j (at most 2^24)
            T* const blk = v + blocks[bfirst + j];
            for (int r = threadIdx.x; r < cnt; r += blockDim.x) {
                const int lo = r & ((1 << B) - 1);
                const int ik = (r >> B) & ((1 << j) - 1);
                const int hi = r >> (B + j);
                T left = T(0), right = T(0);
                if (ik != 0) {                      // left parent, asg.c:828-840
                    const int tz = __builtin_ctz(ik);
                    const int pl = j - 1 - tz;
                    const int pi = ik >> (tz + 1);
                    left = v[blocks[bfirst + pl] + (hi << (B + pl)) + (pi << B) + lo];
                }
                if (ik != (1 << j) - 1) {           // right parent, asg.c:842-854
                    const int tz = __builtin_ctz(ik + 1);
                    const int pl = j - 1 - tz;
                    const int pi = (ik + 1) >> (tz + 1);
                    right = v[blocks[bfirst + pl] + (hi << (B + pl)) + (pi << B) + lo];
                }
                blk[r] = blk[r] - (left + right) * T(0.5);
            }
            __syncthreads();
        }
    }
}

// One workgroup per tile of plane families of the dimension pair (k, k+1): the tile is
// loaded into LDS, dimension k is swept (every value from the values as loaded), then
// dimension k+1 (from the values after the k sweep), and the tile is written back once.
// Each thread keeps the new values of its PER points in registers between reading and
// writing, so every parent is read before its own update, as in asg.c:557-571.  The
// tile's family and sub-pole tables sit in LDS.  Point layout: see PairTables.
template <class T, int PER>
__global__ __launch_bounds__(kThreads) void k_hierarchize_pair(T* __restrict__ v,
                                                               const int32_t* __restrict__ tiles,
                                                               const int32_t* __restrict__ fams,
                                                               const int32_t* __restrict__ subs,
                                                               const int32_t* __restrict__ blocks,
                                                               int lds_values, int lds_fams,
                                                               int lds_subs) {
    extern __shared__ __align__(16) unsigned char smem[];
    T* buf = reinterpret_cast<T*>(smem);
    int32_t* fst = reinterpret_cast<int32_t*>(smem + size_t(lds_values) * sizeof(T));
    int32_t* fmeta = fst + lds_fams;
    int32_t* fsub = fmeta + lds_fams;
    int32_t* sfirst = fsub + lds_fams;          // sub-pole starts within their family
    int32_t* sblock = sfirst + lds_subs;        // sub-pole first block, within the tile
    int32_t* sblk = sblock + lds_subs;          // the tile's block bases
    // the tile's entry names its families, sub-poles and blocks (each consecutive), so
    // the three tables load side by side
    const int32_t* te = tiles + kPairTileInts * blockIdx.x;
    const int f0 = te[0], nf = te[1], nv = te[2];
    const int s0 = te[3], nsub = te[4], b0 = te[5], nb = te[6];
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
        fst[i] = fams[3 * (f0 + i)];
        fmeta[i] = fams[3 * (f0 + i) + 1];
        fsub[i] = fams[3 * (f0 + i) + 2] - s0;
    }
    for (int i = threadIdx.x; i < nsub; i += blockDim.x) {
        sfirst[i] = subs[2 * (s0 + i)];
        sblock[i] = subs[2 * (s0 + i) + 1] - b0;
    }
    for (int i = threadIdx.x; i < nb; i += blockDim.x) sblk[i] = blocks[b0 + i];
    __syncthreads();

    int pf[PER], pa[PER], pb[PER], pr[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int val = threadIdx.x + i * kThreads;
        pf[i] = pa[i] = pb[i] = pr[i] = 0;
        if (val < nv) {
            int lo = 0, hi = nf - 1;            // last family whose start is <= val
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (fst[mid] <= val) lo = mid;
                else hi = mid - 1;
            }
            const int meta = fmeta[lo];
            const int A = meta >> 16;
            const int sf = fsub[lo];
            const int local = val - fst[lo];
            int a = 0;
            while (a < A && sfirst[sf + a + 1] <= local) ++a;
            const int sl = local - sfirst[sf + a];
            const int cc = (meta & 255) + a;
            const int b = 31 - __clz((sl >> cc) + 1);
            pf[i] = lo;
            pa[i] = a;
            pb[i] = b;
            pr[i] = sl - (((1 << b) - 1) << cc);
            buf[val] = v[sblk[sblock[sf + a] + b] + pr[i]];
        }
    }
    __syncthreads();

    T nv_[PER];
    // dimension k: parent (pl, b) in sub-pole pl, in-block index hi | pi | i_{k+1} lo
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int val = threadIdx.x + i * kThreads;
        nv_[i] = T(0);
        if (val < nv) {
            T x = buf[val];
            const int a = pa[i], b = pb[i], r = pr[i];
            if (a > 0) {
                const int meta = fmeta[pf[i]];
                const int c = meta & 255, Bfix = (meta >> 8) & 255;
                const int fb = fst[pf[i]], sf = fsub[pf[i]];
                const int sh = b + Bfix;
                const int midlo = r & ((1 << sh) - 1);
                const int ik = (r >> sh) & ((1 << a) - 1);
                const int hi = r >> (sh + a);
                T left = T(0), right = T(0);
                if (ik != 0) {                              // asg.c:828-840
                    const int tz = __builtin_ctz(ik);
                    const int pl = a - 1 - tz;
                    const int pi = ik >> (tz + 1);
                    left = buf[fb + sfirst[sf + pl] + (((1 << b) - 1) << (c + pl)) +
                               (hi << (sh + pl)) + (pi << sh) + midlo];
                }
                if (ik != (1 << a) - 1) {                   // asg.c:842-854
                    const int tz = __builtin_ctz(ik + 1);
                    const int pl = a - 1 - tz;
                    const int pi = (ik + 1) >> (tz + 1);
                    right = buf[fb + sfirst[sf + pl] + (((1 << b) - 1) << (c + pl)) +
                                (hi << (sh + pl)) + (pi << sh) + midlo];
                }
                x = x - (left + right) * T(0.5);
            }
            nv_[i] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int val = threadIdx.x + i * kThreads;
        if (val < nv) buf[val] = nv_[i];
    }
    __syncthreads();
    // dimension k+1: within the sub-pole, as the one-dimension sweep
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int val = threadIdx.x + i * kThreads;
        if (val < nv) {
            T x = buf[val];
            const int a = pa[i], b = pb[i], r = pr[i];
            const int sf = fsub[pf[i]];
            if (b > 0) {
                const int meta = fmeta[pf[i]];
                const int Bfix = (meta >> 8) & 255;
                const int cc = (meta & 255) + a;
                const int base = fst[pf[i]] + sfirst[sf + a];
                const int lo = r & ((1 << Bfix) - 1);
                const int ik = (r >> Bfix) & ((1 << b) - 1);
                const int hi = r >> (Bfix + b);
                T left = T(0), right = T(0);
                if (ik != 0) {
                    const int tz = __builtin_ctz(ik);
                    const int pl = b - 1 - tz;
                    const int pi = ik >> (tz + 1);
                    left = buf[base + (((1 << pl) - 1) << cc) + (hi << (Bfix + pl)) +
                               (pi << Bfix) + lo];
                }
                if (ik != (1 << b) - 1) {
                    const int tz = __builtin_ctz(ik + 1);
                    const int pl = b - 1 - tz;
                    const int pi = (ik + 1) >> (tz + 1);
                    right = buf[base + (((1 << pl) - 1) << cc) + (hi << (Bfix + pl)) +
                                (pi << Bfix) + lo];
                }
                x = x - (left + right) * T(0.5);
            }
            v[sblk[sblock[sf + a] + b] + r] = x;
        }
    }
}

}  // namespace

template <class T>
int launch_hierarchize_pairs(T* v, const int32_t* tiles, const int32_t* fams, const int32_t* subs,
                             const int32_t* blocks, int num_tiles, int tile_values,
                             int max_families, int max_subs, int max_blocks, hipStream_t s) {
    if (tile_values > kPairMaxValues) {
        set_error("hierarchize: plane-family tile larger than the kernel's limit");
        return kUnsupported;
    }
    // LDS: the largest tile's values, family, sub-pole and block tables
    const size_t lds = size_t(tile_values) * sizeof(T) + size_t(max_families) * 12 +
                       size_t(max_subs) * 8 + size_t(max_blocks) * 4;
    const int per = (tile_values + kThreads - 1) / kThreads;
#define SG_PAIR(PER)                                                                        \
    hipLaunchKernelGGL((k_hierarchize_pair<T, PER>), dim3(num_tiles), dim3(kThreads), lds, s, \
                       v, tiles, fams, subs, blocks, tile_values, max_families, max_subs)
    if (per <= 2) SG_PAIR(2);
    else if (per <= 4) SG_PAIR(4);
    else if (per <= 8) SG_PAIR(8);
    else SG_PAIR(16);
#undef SG_PAIR
    return check_launch("k_hierarchize_pair");
}
template int launch_hierarchize_pairs<double>(double*, const int32_t*, const int32_t*,
                                              const int32_t*, const int32_t*, int, int, int, int,
                                              int, hipStream_t);
template int launch_hierarchize_pairs<float>(float*, const int32_t*, const int32_t*,
                                             const int32_t*, const int32_t*, int, int, int, int,
                                             int, hipStream_t);

namespace {

template <class T>
int hierarchize_impl(const Grid& g, const DeviceTables& t, T* v, hipStream_t s) {
    for (int k = 0; k < g.d;) {
        const bool pair = (k % 2 == 0) && k + 1 < g.d && t.pairs[k / 2].ok;
        if (pair) {
            const DeviceTables::Pair& pr = t.pairs[k / 2];
            if (pr.num_tiles > 0) {
                int st = launch_hierarchize_pairs<T>(v, pr.tiles, pr.fams, pr.subs, pr.blocks,
                                                     pr.num_tiles, pr.tile_values,
                                                     pr.max_families, pr.max_subs,
                                                     pr.max_blocks, s);
                if (st != kOk) return st;
            }
            k += 2;
            continue;
        }
        const DeviceTables::Dim& dm = t.dims[k];
        if (dm.num_tiles > 0) {
            int st = launch_hierarchize_tiles<T>(v, dm.tiles, dm.fams, dm.blocks, dm.num_tiles,
                                                 dm.tile_values, dm.max_families, s);
            if (st != kOk) return st;
        }
        ++k;
    }
    return kOk;
}

// -------------------------------------------------------------------- evaluate --
// The depth-first walk kernels live in sg_eval.hpp and are instantiated in
// sg_eval_f64.hip / sg_eval_f32.hip.

// Reference operation order (asg.c:1250-1268): for each subspace in compact order the hat
// product over k = 0..d-1 starting from 1, times the surplus, added to the running sum.
// Division by div = 2^-l is exact, so it is written as ldexp.  No contraction.
template <int D, class T>
__global__ __launch_bounds__(kThreads) void k_eval_ref(const T* __restrict__ alpha,
                                                       const T* __restrict__ coords,
                                                       T* __restrict__ out, int64_t P,
                                                       const uint8_t* __restrict__ sub_levels,
                                                       const int32_t* __restrict__ sub_base,
                                                       int64_t S) {
#pragma clang fp contract(off)
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = p < P;
    const T* xp = coords + (live ? p : 0) * D;
    T x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = xp[k];
    T acc = T(0);
    for (int64_t j = 0; j < S; ++j) {
        const uint8_t* lv = sub_levels + j * D;
        T prod = T(1);
        int32_t index2 = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const int l = lv[k];
            const T div = ldexp(T(1), -l);
            const int t = static_cast<int>(ldexp(x[k], l));
            index2 = index2 * (1 << l) + t;
            const T left = T(t) * div;
            const T m = ldexp(T(2) * (x[k] - left) - div, l);
            prod *= T(1) + m * (T(1) - T((m >= T(0)) << 1));
        }
        prod *= alpha[sub_base[j] + index2];
        acc += prod;
    }
    if (live) out[p] = acc;
}

template <class T>
int evaluate_ref_impl(const Grid& g, const DeviceTables& t, const T* alpha, const T* coords,
                      int64_t P, T* out, hipStream_t s) {
    if (P == 0) return kOk;
    const int64_t blocks = (P + kThreads - 1) / kThreads;
    if (blocks > int64_t(0x7fffffff)) {
        set_error("evaluate_ref: too many points for one launch");
        return kRange;
    }
    switch (g.d) {
#define SG_CASE(DD)                                                                        \
    case DD:                                                                               \
        hipLaunchKernelGGL((k_eval_ref<DD, T>), dim3(unsigned(blocks)), dim3(kThreads), 0, \
                           s, alpha, coords, out, P, t.sub_levels, t.sub_base, g.S);       \
        return check_launch("k_eval_ref");
        SG_CASE(1) SG_CASE(2) SG_CASE(3) SG_CASE(4) SG_CASE(5) SG_CASE(6) SG_CASE(7)
        SG_CASE(8) SG_CASE(9) SG_CASE(10) SG_CASE(11) SG_CASE(12) SG_CASE(13) SG_CASE(14)
        SG_CASE(15) SG_CASE(16) SG_CASE(17) SG_CASE(18) SG_CASE(19) SG_CASE(20)
#undef SG_CASE
        default:
            set_error("evaluate_ref: unsupported number of dimensions");
            return kUnsupported;
    }
}

// -------------------------------------------------------------- point binning --

// key = interleaved leading bits of the coordinates: bit 1 of dims 0..d-1, then bit 2,
// ...; the histogram of keys is accumulated in the same pass.
template <class T>
__global__ __launch_bounds__(kThreads) void k_point_keys(const T* __restrict__ coords, int64_t P,
                                                         int d, int key_bits,
                                                         uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ hist) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P;
         i += int64_t(gridDim.x) * blockDim.x) {
        const T* x = coords + i * d;
        uint32_t key = 0;
        const int levels = (key_bits + d - 1) / d;
        for (int j = 0; j < levels; ++j) {
            for (int k = 0; k < d; ++k) {
                if (j * d + k >= key_bits) break;
                const uint32_t t = static_cast<uint32_t>(ldexp(double(x[k]), j + 1));
                key = (key << 1) | (t & 1u);
            }
        }
        keys[i] = key;
        atomicAdd(&hist[key], 1u);
    }
}

// exclusive scan of n counts in place, one workgroup of kThreads threads
__global__ __launch_bounds__(kThreads) void k_scan_counts(uint32_t* __restrict__ c, int64_t n) {
    __shared__ uint32_t part[kThreads];
    const int64_t chunk = (n + kThreads - 1) / kThreads;
    const int64_t lo = threadIdx.x * chunk;
    const int64_t hi = lo + chunk < n ? lo + chunk : n;
    uint32_t sum = 0;
    for (int64_t i = lo; i < hi; ++i) sum += c[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < kThreads; ++i) {
            const uint32_t v = part[i];
            part[i] = run;
            run += v;
        }
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
    for (int64_t i = lo; i < hi; ++i) {
        const uint32_t v = c[i];
        c[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kThreads) void k_scatter_points(const uint32_t* __restrict__ keys,
                                                             int64_t P,
                                                             uint32_t* __restrict__ offs,
                                                             uint32_t* __restrict__ perm) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < P;
         i += int64_t(gridDim.x) * blockDim.x)
        perm[atomicAdd(&offs[keys[i]], 1u)] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(kThreads) void k_fill_u32(uint32_t* __restrict__ p, int64_t n,
                                                       uint32_t v) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

template <class T>
int sort_points_impl(const T* coords, int64_t P, int d, int key_bits, void* ws, uint32_t* perm,
                     hipStream_t s) {
    if (P == 0) return kOk;
    if (P > int64_t(0xffffffff) || key_bits < 0 || key_bits > 24 || d < 1) {
        set_error("sort_points: bad size");
        return kInvalid;
    }
    uint32_t* keys = static_cast<uint32_t*>(ws);
    uint32_t* hist = keys + P;
    const int64_t nb = int64_t(1) << key_bits;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nb)), dim3(kThreads), 0, s, hist, nb, 0u);
    hipLaunchKernelGGL(k_point_keys<T>, dim3(grid_for(P)), dim3(kThreads), 0, s, coords, P, d,
                       key_bits, keys, hist);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kThreads), 0, s, hist, nb);
    hipLaunchKernelGGL(k_scatter_points, dim3(grid_for(P)), dim3(kThreads), 0, s, keys, P, hist,
                       perm);
    return check_launch("sort_points");
}

// ------------------------------------------------------------------ utilities --

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_uniform(int64_t count, uint64_t seed, T* out) {
    for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < count;
         j += int64_t(gridDim.x) * blockDim.x) {
        const uint64_t k = mix64((seed << 32) + uint64_t(j)) >> 40;
        out[j] = T(ldexp(double(k), -24));
    }
}

// The achievable-gather pattern (see launch_gather_probe_f64): lane p draws an xorshift32
// stream and gathers element (r & (2^s - 1)) of each subspace block, 8 at a time, with the
// same raw buffer loads (block base in the scalar offset) as the evaluation walk.
template <class T>
__global__ __launch_bounds__(kThreads) void k_gather_probe(const T* alpha, int64_t N,
                                                           const int32_t* base,
                                                           const uint8_t* lsum, int S,
                                                           int64_t P, uint32_t seed, T* out) {
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint32_t r = uint32_t(mix64(uint64_t(seed) << 32 | uint64_t(p))) | 1u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T*>(alpha), short(0), int(N * int64_t(sizeof(T))), 0x00020000);
    const auto* cb = (const __attribute__((address_space(4))) int32_t*)base;
    const auto* cl = (const __attribute__((address_space(4))) uint8_t*)lsum;
    T acc = T(0);
    for (int i = 0; i < S; i += 8) {
        T v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            v[j] = T(0);
            if (i + j < S) {
                r ^= r << 13;
                r ^= r >> 17;
                r ^= r << 5;
                const uint32_t off = (r & ((1u << cl[i + j]) - 1u)) * uint32_t(sizeof(T));
                const uint32_t so = uint32_t(cb[i + j]) * uint32_t(sizeof(T));
                if constexpr (sizeof(T) == 8)
                    v[j] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, off, so, 0));
                else
                    v[j] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, off, so, 0));
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) acc += v[j];
    }
    if (p < P) out[p] = acc;
}

// out[p][k] = the point p, coordinate k of a column-major or 32-point-tiled input; the
// writes are coalesced, the reads strided (a one-pass transpose at memory speed).
template <class T>
__global__ __launch_bounds__(kThreads) void k_points_to_row_major(const T* __restrict__ in,
                                                                  int layout, int d, int64_t P,
                                                                  T* __restrict__ out) {
    const int64_t total = P * d;
    for (int64_t o = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; o < total;
         o += int64_t(gridDim.x) * blockDim.x) {
        const int64_t p = o / d;
        const int k = int(o - p * d);
        int64_t src;
        if (layout == kPointsColumnMajor) src = int64_t(k) * P + p;
        else if (layout == kPointsTiled32) src = (p >> 5) * d * 32 + int64_t(k) * 32 + (p & 31);
        else src = o;
        out[o] = in[src];
    }
}

__global__ __launch_bounds__(kThreads) void k_f64_to_f32(const double* in, float* out,
                                                         int64_t count) {
    for (int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < count;
         j += int64_t(gridDim.x) * blockDim.x)
        out[j] = float(in[j]);
}

template <class T>
void free_dev(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <class T>
int upload(T*& dst, const T* src, size_t count) {
    if (count == 0) return kOk;
    hipError_t e = hipMalloc(&dst, count * sizeof(T));
    if (e != hipSuccess) {
        dst = nullptr;
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? kNoMem : kHip;
    }
    SG_HIP_CHECK(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return kOk;
}

}  // namespace

template <class T>
int launch_hierarchize_tiles(T* v, const int32_t* tiles, const int32_t* fams,
                             const int32_t* blocks, int num_tiles, int tile_values,
                             int max_families, hipStream_t s) {
    const size_t lds = size_t(tile_values) * sizeof(T) + size_t(max_families) * 4;
    if (lds > kPoleLdsBytes) {
        hipLaunchKernelGGL(k_hierarchize_dim_global<T>, dim3(num_tiles), dim3(kThreads), 0, s, v,
                           tiles, fams, blocks);
        return check_launch("k_hierarchize_dim_global");
    }
    hipLaunchKernelGGL(k_hierarchize_dim<T>, dim3(num_tiles), dim3(kThreads), lds, s, v, tiles,
                       fams, blocks, tile_values);
    return check_launch("k_hierarchize_dim");
}
template int launch_hierarchize_tiles<double>(double*, const int32_t*, const int32_t*,
                                              const int32_t*, int, int, int, hipStream_t);
template int launch_hierarchize_tiles<float>(float*, const int32_t*, const int32_t*,
                                             const int32_t*, int, int, int, hipStream_t);

// ------------------------------------------------------------------- launchers --

int upload_tables(const Grid& g, DeviceTables& t) {
    SG_HIP_CHECK(hipGetDevice(&t.device));
    std::vector<int32_t> dp32(g.dp.size()), lidx32(g.lidx.size()), cnt32(g.cnt.size());
    for (size_t i = 0; i < g.dp.size(); ++i)
        dp32[i] = static_cast<int32_t>(std::min<int64_t>(g.dp[i], 0x7fffffff));
    for (size_t i = 0; i < g.lidx.size(); ++i) lidx32[i] = static_cast<int32_t>(g.lidx[i]);
    for (size_t i = 0; i < g.cnt.size(); ++i) cnt32[i] = static_cast<int32_t>(g.cnt[i]);
    int st;
    if ((st = upload(t.dp, dp32.data(), dp32.size())) != kOk) return st;
    if ((st = upload(t.lidx, lidx32.data(), lidx32.size())) != kOk) return st;
    if ((st = upload(t.cnt, cnt32.data(), cnt32.size())) != kOk) return st;
    if ((st = upload(t.sub_levels, g.sub_levels.data(), g.sub_levels.size())) != kOk) return st;
    if ((st = upload(t.sub_base, g.sub_base.data(), g.sub_base.size())) != kOk) return st;
    {
        std::vector<uint8_t> ls(size_t(g.S), 0);
        for (int64_t i = 0; i < g.S; ++i)
            for (int j = 0; j < g.d; ++j) ls[size_t(i)] = uint8_t(ls[size_t(i)] + g.sub_levels[size_t(i * g.d + j)]);
        if ((st = upload(t.sub_lsum, ls.data(), ls.size())) != kOk) return st;
    }
    if ((st = upload(t.dfs_base, g.dfs_base.data(), g.dfs_base.size())) != kOk) return st;
    if ((st = upload(t.dfs_bound, g.dfs_bound.data(), g.dfs_bound.size())) != kOk) return st;
    {
        // staged-pattern probe tables from the walks' stage plans (each plan lists its
        // leaves in its own walk order)
        for (int kind = 0; kind < 2; ++kind)
            for (int e = 0; e < 2; ++e) {
                const bool ok = kind == 0 ? g.stage[e].ok : g.blk[e].ok;
                if (!ok) continue;
                const std::vector<int32_t>& bases = kind == 0 ? g.dfs_base : g.blk[e].leaf_base;
                std::vector<int32_t> lsum(size_t(g.S), 0);
                for (int64_t q = 0; q < g.S; ++q) {
                    const auto it =
                        std::lower_bound(g.sub_base.begin(), g.sub_base.end(), bases[size_t(q)]);
                    const int64_t id = it - g.sub_base.begin();
                    int sum = 0;
                    for (int j = 0; j < g.d; ++j) sum += g.sub_levels[size_t(id * g.d + j)];
                    lsum[size_t(q)] = sum;
                }
                const std::vector<int32_t>& lds = kind == 0 ? g.stage[e].leaf_lds : g.blk[e].leaf_lds;
                const std::vector<int32_t>& first =
                    kind == 0 ? g.stage[e].stage_first : g.blk[e].stage_first;
                const int elem = e == 0 ? 8 : 4;
                const int K = int(first.size()) - 1;
                std::vector<int32_t> info, glob;
                for (int k = 0; k < K; ++k) {
                    int32_t cnt = 0, lo = INT32_MAX, hi = 0;
                    for (int32_t q = first[k]; q < first[k + 1]; ++q) {
                        if (lds[size_t(q)] < 0) continue;
                        ++cnt;
                        lo = std::min(lo, lds[size_t(q)]);
                        hi = std::max(hi, lds[size_t(q)] + (int32_t(1) << lsum[size_t(q)]) * elem);
                    }
                    int64_t cells = cnt ? (hi - lo) / elem : 1, pw = 1;
                    while (pw * 2 <= cells) pw *= 2;     // largest power of two inside the stage
                    info.push_back(cnt);
                    info.push_back(int32_t(pw - 1));
                }
                for (int64_t q = 0; q < g.S; ++q)
                    if (lds[size_t(q)] < 0) {
                        glob.push_back(bases[size_t(q)]);
                        glob.push_back(lsum[size_t(q)]);
                    }
                DeviceTables::Probe& pr = t.probe[kind][e];
                pr.num_stages = K;
                pr.num_glob = int(glob.size() / 2);
                if (info.empty()) info.push_back(0);
                if (glob.empty()) glob.push_back(0);
                if ((st = upload(pr.info, info.data(), info.size())) != kOk) return st;
                if ((st = upload(pr.glob, glob.data(), glob.size())) != kOk) return st;
            }
    }
    eval::GridDesc desc{};
    desc.leaf_base = t.dfs_base;
    desc.bounds = t.dfs_bound;
    desc.N = g.N;
    desc.n = g.n;
    for (int j = 0; j < g.d; ++j) desc.limits[j] = g.limits[j];
    for (int e = 0; e < 2; ++e) {
        const StagePlan& p = g.stage[e];
        if (!p.ok) continue;
        // one allocation: leaf_lds | stage_first | seg_begin | segs
        std::vector<int32_t> all;
        all.reserve(p.leaf_lds.size() + p.stage_first.size() + p.seg_begin.size() + p.segs.size());
        all.insert(all.end(), p.leaf_lds.begin(), p.leaf_lds.end());
        all.insert(all.end(), p.stage_first.begin(), p.stage_first.end());
        all.insert(all.end(), p.seg_begin.begin(), p.seg_begin.end());
        all.insert(all.end(), p.segs.begin(), p.segs.end());
        if ((st = upload(t.stage_tabs[e], all.data(), all.size())) != kOk) return st;
        eval::GridDesc::Stages& sd = desc.stage[e];
        sd.leaf_lds = t.stage_tabs[e];
        sd.stage_first = sd.leaf_lds + p.leaf_lds.size();
        sd.seg_begin = sd.stage_first + p.stage_first.size();
        sd.segs = sd.seg_begin + p.seg_begin.size();
        sd.num_stages = int32_t(p.stage_first.size()) - 1;
    }
    for (int e = 0; e < 2; ++e) {
        const BlockStagePlan& p = g.blk[e];
        if (!p.ok) continue;
        std::vector<int32_t> all(p.seg_begin);
        all.insert(all.end(), p.segs.begin(), p.segs.end());
        all.insert(all.end(), p.leaf_base.begin(), p.leaf_base.end());
        if ((st = upload(t.blk_tabs[e], all.data(), all.size())) != kOk) return st;
        eval::GridDesc::BlkStages& sd = desc.blk[e];
        sd.seg_begin = t.blk_tabs[e];
        sd.segs = sd.seg_begin + p.seg_begin.size();
        sd.leaf_base = sd.segs + p.segs.size();
        sd.num_stages = p.num_stages;
    }
    t.host_desc = desc;
    eval::GridDesc* ddesc = nullptr;
    if ((st = upload(ddesc, &desc, 1)) != kOk) return st;
    t.desc = ddesc;
    for (int k = 0; k < g.d; ++k) {
        const HierTables& h = g.hier[k];
        DeviceTables::Dim& dm = t.dims[k];
        dm.num_tiles = int(h.tiles.size() / 3);
        dm.max_families = h.max_families_per_tile;
        dm.tile_values = 0;
        for (int i = 0; i < dm.num_tiles; ++i) dm.tile_values = std::max(dm.tile_values, h.tiles[3 * i + 2]);
        if ((st = upload(dm.tiles, h.tiles.data(), h.tiles.size())) != kOk) return st;
        if ((st = upload(dm.fams, h.fams.data(), h.fams.size())) != kOk) return st;
        if ((st = upload(dm.blocks, h.blocks.data(), h.blocks.size())) != kOk) return st;
    }
    for (size_t p = 0; p < g.hier2.size(); ++p) {
        const PairTables& h = g.hier2[p];
        DeviceTables::Pair& pr = t.pairs[p];
        pr.ok = h.ok;
        if (!h.ok) continue;
        pr.num_tiles = int(h.tiles.size() / kPairTileInts);
        pr.max_families = h.max_families_per_tile;
        pr.max_subs = h.max_subs_per_tile;
        pr.max_blocks = h.max_blocks_per_tile;
        pr.tile_values = h.max_tile_values;
        if ((st = upload(pr.tiles, h.tiles.data(), h.tiles.size())) != kOk) return st;
        if ((st = upload(pr.fams, h.fams.data(), h.fams.size())) != kOk) return st;
        if ((st = upload(pr.subs, h.subs.data(), h.subs.size())) != kOk) return st;
        if ((st = upload(pr.blocks, h.blocks.data(), h.blocks.size())) != kOk) return st;
    }
    return kOk;
}

void free_tables(DeviceTables& t) {
    int cur = -1;
    bool switched = false;
    if (t.device >= 0 && hipGetDevice(&cur) == hipSuccess && cur != t.device)
        switched = hipSetDevice(t.device) == hipSuccess;
    free_dev(t.dp);
    free_dev(t.lidx);
    free_dev(t.cnt);
    free_dev(t.sub_levels);
    free_dev(t.sub_base);
    free_dev(t.sub_lsum);
    free_dev(t.dfs_base);
    free_dev(t.dfs_bound);
    free_dev(t.stage_tabs[0]);
    free_dev(t.stage_tabs[1]);
    free_dev(t.blk_tabs[0]);
    free_dev(t.blk_tabs[1]);
    for (auto& pk : t.probe)
        for (auto& pr : pk) {
            free_dev(pr.info);
            free_dev(pr.glob);
        }
    if (t.desc) (void)hipFree(t.desc);
    t.desc = nullptr;
    for (auto& dm : t.dims) {
        free_dev(dm.tiles);
        free_dev(dm.fams);
        free_dev(dm.blocks);
        dm.num_tiles = 0;
    }
    for (auto& pr : t.pairs) {
        free_dev(pr.tiles);
        free_dev(pr.fams);
        free_dev(pr.subs);
        free_dev(pr.blocks);
        pr.num_tiles = 0;
        pr.ok = false;
    }
    if (switched) (void)hipSetDevice(cur);
}

int launch_idx2gp(const Grid& g, const DeviceTables& t, const int32_t* index, int64_t count,
                  int32_t* levels, int32_t* indices, hipStream_t s) {
    if (count == 0) return kOk;
    hipLaunchKernelGGL(k_idx2gp, dim3(grid_for(count)), dim3(kThreads), 0, s, make_bij(g, t),
                       index, count, levels, indices);
    return check_launch("k_idx2gp");
}

int launch_gp2idx(const Grid& g, const DeviceTables& t, const int32_t* levels,
                  const int32_t* indices, int64_t count, int32_t* index, hipStream_t s) {
    if (count == 0) return kOk;
    hipLaunchKernelGGL(k_gp2idx, dim3(grid_for(count)), dim3(kThreads), 0, s, make_bij(g, t),
                       levels, indices, count, index);
    return check_launch("k_gp2idx");
}

int launch_grid_coords_f64(const Grid& g, const DeviceTables& t, double* coords, hipStream_t s) {
    hipLaunchKernelGGL(k_grid_coords, dim3(grid_for(g.N)), dim3(kThreads), 0, s, make_bij(g, t),
                       coords);
    return check_launch("k_grid_coords");
}

int launch_hierarchize_f64(const Grid& g, const DeviceTables& t, double* v, hipStream_t s) {
    return hierarchize_impl<double>(g, t, v, s);
}
int launch_hierarchize_f32(const Grid& g, const DeviceTables& t, float* v, hipStream_t s) {
    return hierarchize_impl<float>(g, t, v, s);
}

int launch_evaluate_ref_f64(const Grid& g, const DeviceTables& t, const double* alpha,
                            const double* coords, int64_t P, double* out, hipStream_t s) {
    return evaluate_ref_impl<double>(g, t, alpha, coords, P, out, s);
}
int launch_evaluate_ref_f32(const Grid& g, const DeviceTables& t, const float* alpha,
                            const float* coords, int64_t P, float* out, hipStream_t s) {
    return evaluate_ref_impl<float>(g, t, alpha, coords, P, out, s);
}

int key_bits_for(int64_t P, int d) {
    // about 32 points per key (a wave of 64 lanes spans two or three keys), at most 2^16
    // keys so the one-workgroup scan stays short
    int b = 0;
    while (b < 16 && (int64_t(64) << b) <= P) ++b;
    return std::min(b, 20 * d);
}

size_t sort_workspace_bytes(int64_t P, int key_bits) {
    return size_t(P) * 4 + ((size_t(1) << key_bits) + 1) * 4;
}

// Scratch of the allocation-free evaluate entry point: the binning workspace (keys and
// key histogram) followed by the point permutation, for up to kNoAllocMaxPoints points.
// Every device holds its own copy of the array for the life of the process.  Host code
// uses the address hipGetSymbolAddress returns on that device, never the array's name
// (which is a host-side stand-in).
constexpr size_t kNoAllocScratchBytes =
    (size_t(kNoAllocMaxPoints) * 4 + ((size_t(1) << 16) + 1) * 4 + size_t(kNoAllocMaxPoints) * 4 +
     255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

int noalloc_scratch(void** ptr, size_t* bytes) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = -1;
    SG_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("noalloc scratch: device ordinal out of range");
        return kInvalid;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        SG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("noalloc scratch: hipGetSymbolAddress returned no address");
            return kHip;
        }
        addr[dev] = p;
    }
    *ptr = addr[dev];
    *bytes = kNoAllocScratchBytes;
    return kOk;
}

int launch_sort_points_f64(const double* coords, int64_t P, int d, int key_bits, void* ws,
                           uint32_t* perm, hipStream_t s) {
    return sort_points_impl<double>(coords, P, d, key_bits, ws, perm, s);
}
int launch_sort_points_f32(const float* coords, int64_t P, int d, int key_bits, void* ws,
                           uint32_t* perm, hipStream_t s) {
    return sort_points_impl<float>(coords, P, d, key_bits, ws, perm, s);
}

int launch_uniform_points_f64(int64_t count, uint64_t seed, double* out, hipStream_t s) {
    if (count == 0) return kOk;
    hipLaunchKernelGGL(k_uniform<double>, dim3(grid_for(count)), dim3(kThreads), 0, s, count,
                       seed, out);
    return check_launch("k_uniform");
}
int launch_uniform_points_f32(int64_t count, uint64_t seed, float* out, hipStream_t s) {
    if (count == 0) return kOk;
    hipLaunchKernelGGL(k_uniform<float>, dim3(grid_for(count)), dim3(kThreads), 0, s, count,
                       seed, out);
    return check_launch("k_uniform");
}
template <class T>
static int gather_probe(const Grid& g, const DeviceTables& t, const T* alpha, int64_t P,
                        uint32_t seed, T* out, hipStream_t s) {
    if (P <= 0) return kOk;
    if (g.N * int64_t(sizeof(T)) > int64_t(0x7fffffff) || (P + kThreads - 1) / kThreads > 0x7fffffff) {
        set_error("gather probe: grid or point count out of range");
        return kRange;
    }
    hipLaunchKernelGGL(k_gather_probe<T>, dim3(unsigned((P + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, alpha, g.N, t.sub_base, t.sub_lsum, int(g.S), P,
                       seed, out);
    return check_launch("k_gather_probe");
}
int launch_gather_probe_f64(const Grid& g, const DeviceTables& t, const double* alpha,
                            int64_t P, uint32_t seed, double* out, hipStream_t s) {
    return gather_probe(g, t, alpha, P, seed, out, s);
}
int launch_gather_probe_f32(const Grid& g, const DeviceTables& t, const float* alpha,
                            int64_t P, uint32_t seed, float* out, hipStream_t s) {
    return gather_probe(g, t, alpha, P, seed, out, s);
}

template <class T>
static int sample(const Grid& g, const DeviceTables& t, const SampleParams& sp, T* out,
                  hipStream_t s) {
    if (sp.fn < kFnReference || sp.fn > kFnSinProduct || sp.terms < 1 ||
        sp.terms > kMaxSampleTerms) {
        set_error("sample: unknown function or term count");
        return kInvalid;
    }
    hipLaunchKernelGGL(k_sample<T>, dim3(grid_for(g.N)), dim3(kThreads), 0, s, make_bij(g, t),
                       sp, out);
    return check_launch("k_sample");
}
int launch_sample_f64(const Grid& g, const DeviceTables& t, const SampleParams& sp, double* out,
                      hipStream_t s) {
    return sample(g, t, sp, out, s);
}
int launch_sample_f32(const Grid& g, const DeviceTables& t, const SampleParams& sp, float* out,
                      hipStream_t s) {
    return sample(g, t, sp, out, s);
}

template <class T>
static int points_to_row_major(const T* in, int layout, int d, int64_t P, T* out, hipStream_t s) {
    if (layout < kPointsRowMajor || layout > kPointsTiled32 || d < 1 || d > kMaxDims || P < 0) {
        set_error("points_to_row_major: bad layout, dimension or count");
        return kInvalid;
    }
    if (P == 0) return kOk;
    hipLaunchKernelGGL(k_points_to_row_major<T>, dim3(grid_for(P * d)), dim3(kThreads), 0, s, in,
                       layout, d, P, out);
    return check_launch("k_points_to_row_major");
}
int launch_points_to_row_major_f64(const double* in, int layout, int d, int64_t P, double* out,
                                   hipStream_t s) {
    return points_to_row_major(in, layout, d, P, out, s);
}
int launch_points_to_row_major_f32(const float* in, int layout, int d, int64_t P, float* out,
                                   hipStream_t s) {
    return points_to_row_major(in, layout, d, P, out, s);
}

int launch_f64_to_f32(const double* in, float* out, int64_t count, hipStream_t s) {
    if (count == 0) return kOk;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(count)), dim3(kThreads), 0, s, in, out, count);
    return check_launch("k_f64_to_f32");
}

}  // namespace sg
