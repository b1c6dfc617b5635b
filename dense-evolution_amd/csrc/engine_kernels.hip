// HIP kernels for gfx950: the LDS-tile stage engine, Pauli-sum kernels and
// FP64 reductions.
//
// The stage kernel replaces the reference's per-gate gather step
// (dense_evolution/circuits/compiler.py:136-315), which rebuilt an int64
// index vector and gathered the whole state once per gate.  Here a workgroup
// loads one tile of 2^m amplitudes with 16-byte contiguous accesses (the low
// `L` index bits are always part of the tile), runs every micro-op of the
// stage on the tile in LDS, and writes it back: one HBM round trip per stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "kernels.h"

namespace desv {

template <typename C> __device__ inline C make_zero();
template <> __device__ inline float2 make_zero<float2>() { return make_float2(0.f, 0.f); }
template <> __device__ inline double2 make_zero<double2>() { return make_double2(0.0, 0.0); }

namespace {

template <typename C> __device__ inline C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename C> __device__ inline C cfma(C a, C b, C acc) {
    acc.x += a.x * b.x - a.y * b.y;
    acc.y += a.x * b.y + a.y * b.x;
    return acc;
}
template <typename C> __device__ inline double2 to_d(C a) { return make_double2((double)a.x, (double)a.y); }

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum `v` over the block; the result is valid in thread 0.  All threads of
// the block must call it.  `red` holds kThreads/64 doubles.
__device__ inline double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    return s;
}

__device__ inline int insert_zero(int i, int p) { return ((i >> p) << (p + 1)) | (i & ((1 << p) - 1)); }

// ---------------------------------------------------------------------------
// Tile micro-ops.  `tile[v]` are NV LDS tiles of 2^m amplitudes.
// ---------------------------------------------------------------------------
template <typename T, int NV>
__device__ inline void op_dense1(typename VT<T>::c **tile, int m, const DevOp &op,
                                 const typename VT<T>::c *M) {
    using C = typename VT<T>::c;
    const C m00 = M[0], m01 = M[1], m10 = M[2], m11 = M[3];
    const int t = op.pos[0];
    const int half = 1 << (m - 1);
    const int cl = (int)op.cloc;
    for (int p = threadIdx.x; p < half; p += blockDim.x) {
        const int i0 = insert_zero(p, t), i1 = i0 | (1 << t);
        if ((i0 & cl) != cl) continue;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            C a0 = tile[v][i0], a1 = tile[v][i1];
            tile[v][i0] = cfma(m01, a1, cmul(m00, a0));
            tile[v][i1] = cfma(m11, a1, cmul(m10, a0));
        }
    }
}

template <typename T, int NV>
__device__ inline void op_dense2(typename VT<T>::c **tile, int m, const DevOp &op,
                                 const typename VT<T>::c *M) {
    using C = typename VT<T>::c;
    C mm[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) mm[i] = M[i];
    const int p0 = op.pos[0], p1 = op.pos[1];
    const int lo = p0 < p1 ? p0 : p1, hi = p0 < p1 ? p1 : p0;
    const int quarter = 1 << (m - 2);
    const int cl = (int)op.cloc;
    for (int q = threadIdx.x; q < quarter; q += blockDim.x) {
        const int b = insert_zero(insert_zero(q, lo), hi);
        if ((b & cl) != cl) continue;
        const int idx[4] = {b, b | (1 << p0), b | (1 << p1), b | (1 << p0) | (1 << p1)};
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            C a[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) a[s] = tile[v][idx[s]];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                C acc = cmul(mm[r * 4 + 0], a[0]);
                acc = cfma(mm[r * 4 + 1], a[1], acc);
                acc = cfma(mm[r * 4 + 2], a[2], acc);
                acc = cfma(mm[r * 4 + 3], a[3], acc);
                tile[v][idx[r]] = acc;
            }
        }
    }
}

template <typename T, int NV, int K>
__device__ inline void op_densek(typename VT<T>::c **tile, int m, const DevOp &op,
                                 const typename VT<T>::c *M) {
    using C = typename VT<T>::c;
    constexpr int D = 1 << K;
    int srt[K];
#pragma unroll
    for (int j = 0; j < K; ++j) srt[j] = op.pos[j];
    // sort ascending for zero insertion
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int b = 0; b < K - 1; ++b)
            if (srt[b] > srt[b + 1]) { int t = srt[b]; srt[b] = srt[b + 1]; srt[b + 1] = t; }
    int off[D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        int o = 0;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (s >> j & 1) o |= 1 << op.pos[j];
        off[s] = o;
    }
    const int groups = 1 << (m - K);
    const int cl = (int)op.cloc;
    for (int g = threadIdx.x; g < groups; g += blockDim.x) {
        int b = g;
#pragma unroll
        for (int j = 0; j < K; ++j) b = insert_zero(b, srt[j]);
        if ((b & cl) != cl) continue;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            C a[D];
#pragma unroll
            for (int s = 0; s < D; ++s) a[s] = tile[v][b | off[s]];
            // rows are not unrolled (keeps code size and compile time small);
            // the row offset is rebuilt from r instead of indexing off[].
#pragma unroll 1
            for (int r = 0; r < D; ++r) {
                C acc = make_zero<C>();
#pragma unroll
                for (int c = 0; c < D; ++c) acc = cfma(M[r * D + c], a[c], acc);
                int o = 0;
#pragma unroll
                for (int j = 0; j < K; ++j)
                    if (r >> j & 1) o |= 1 << op.pos[j];
                tile[v][b | o] = acc;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused k-qubit block (k = 4, 5) as a real-split complex GEMM on MFMA.
//
// The tile is viewed as X[r][g]: r = block-local index over the k target
// bits, g = one of G = 2^(m-k) groups (the other tile bits).  Y = U X with
//   Yr = Ur Xr + Ui (-Xi),   Yi = Ur Xi + Ui Xr,
// computed 16 groups (columns) at a time per wave with 16x16x4 MFMA tiles:
// A fragment  lane l -> U[rb*16 + l%16][ks*4 + l/16]
// B fragment  lane l -> X[ks*4 + l/16][col0 + l%16]
// accumulator lane l, reg j -> Y[rb*16 + 4*(l/16) + j][col0 + l%16].
// A wave reads all of its columns before writing them back, and waves own
// disjoint columns, so the update is in place.
// ---------------------------------------------------------------------------
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
    using acc = f32x4_t;
    static __device__ inline acc op(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<double> {
    using acc = f64x4_t;
    static __device__ inline acc op(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <typename T, int NV, int K>
__device__ inline void op_mfma(typename VT<T>::c **tile, int m, const DevOp &op, const typename VT<T>::c *M) {
    using C = typename VT<T>::c;
    using Acc = typename Mfma<T>::acc;
    constexpr int D = 1 << K;
    constexpr int RB = D / 16;  // 16-row blocks of U
    constexpr int KS = D / 4;   // 4-wide k slices
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, k4 = lane >> 4;
    int pos[K], srt[K];
#pragma unroll
    for (int j = 0; j < K; ++j) pos[j] = srt[j] = op.pos[j];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int b = 0; b < K - 1; ++b)
            if (srt[b] > srt[b + 1]) { int t = srt[b]; srt[b] = srt[b + 1]; srt[b + 1] = t; }
    auto deposit = [&](int r) {
        int o = 0;
#pragma unroll
        for (int j = 0; j < K; ++j) o |= ((r >> j) & 1) << pos[j];
        return o;
    };
    // A fragments (constant over the tile)
    T ar[RB][KS], ai[RB][KS];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const C u = M[(rb * 16 + i16) * D + ks * 4 + k4];
            ar[rb][ks] = u.x;
            ai[rb][ks] = u.y;
        }
    int boff[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) boff[ks] = deposit(ks * 4 + k4);
    int yoff[RB][4];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int j = 0; j < 4; ++j) yoff[rb][j] = deposit(rb * 16 + 4 * k4 + j);

    const int G = 1 << (m - K);
    const int waves = blockDim.x >> 6;
    for (int c0 = (threadIdx.x >> 6) * 16; c0 < G; c0 += waves * 16) {
        const int col = c0 + i16;
        const bool valid = col < G;
        int base = valid ? col : 0;
#pragma unroll
        for (int j = 0; j < K; ++j) base = insert_zero(base, srt[j]);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            T br[KS], bi[KS];
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                C x = tile[v][base | boff[ks]];
                br[ks] = valid ? x.x : T(0);
                bi[ks] = valid ? x.y : T(0);
            }
            Acc yr[RB], yi[RB];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    yr[rb][j] = T(0);
                    yi[rb][j] = T(0);
                }
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    yr[rb] = Mfma<T>::op(ar[rb][ks], br[ks], yr[rb]);
                    yr[rb] = Mfma<T>::op(ai[rb][ks], -bi[ks], yr[rb]);
                    yi[rb] = Mfma<T>::op(ar[rb][ks], bi[ks], yi[rb]);
                    yi[rb] = Mfma<T>::op(ai[rb][ks], br[ks], yi[rb]);
                }
            }
            if (valid) {
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        C y;
                        y.x = yr[rb][j];
                        y.y = yi[rb][j];
                        tile[v][base | yoff[rb][j]] = y;
                    }
            }
        }
    }
}

template <typename T, int NV>
__device__ inline void op_diag(typename VT<T>::c **tile, int m, const DevOp &op,
                               const typename VT<T>::c *Dg, uint64_t base) {
    using C = typename VT<T>::c;
    // Out-of-tile bits are constant over the tile; in-tile bit j of the
    // diagonal index comes from tile bit p[j] (p[j] = 31 marks "not local", a
    // bit that is always 0 for tile indices).  Loops are fully unrolled over
    // the 5 possible bits so nothing is indexed at run time (no scratch).
    int gsub = 0;
    int p[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        p[j] = 31;
        if (j < op.k) {
            const int pj = op.pos[j];
            if (pj < 0) {
                if (base >> (-1 - pj) & 1) gsub |= 1 << j;
            } else {
                p[j] = pj;
            }
        }
    }
    const int size = 1 << m;
    const int cl = (int)op.cloc;
    for (int i = threadIdx.x; i < size; i += blockDim.x) {
        if ((i & cl) != cl) continue;
        int d = gsub;
#pragma unroll
        for (int j = 0; j < 5; ++j) d |= ((i >> p[j]) & 1) << j;
        const C f = Dg[d];
#pragma unroll
        for (int v = 0; v < NV; ++v) tile[v][i] = cmul(f, tile[v][i]);
    }
}

template <typename T, int NV>
__device__ inline void op_scalar(typename VT<T>::c **tile, int m, const typename VT<T>::c *S) {
    using C = typename VT<T>::c;
    const C f = S[0];
    const int size = 1 << m;
    for (int i = threadIdx.x; i < size; i += blockDim.x)
#pragma unroll
        for (int v = 0; v < NV; ++v) tile[v][i] = cmul(f, tile[v][i]);
}

// Re <lam| M |psi> over the tile; tile[0] = psi, tile[1] = lam.
template <typename T>
__device__ inline double op_grad_dense(typename VT<T>::c **tile, int m, const DevOp &op,
                                       const typename VT<T>::c *M) {
    const double2 m00 = to_d(M[0]), m01 = to_d(M[1]), m10 = to_d(M[2]), m11 = to_d(M[3]);
    const int t = op.pos[0];
    const int half = 1 << (m - 1);
    const int cl = (int)op.cloc;
    double acc = 0;
    for (int p = threadIdx.x; p < half; p += blockDim.x) {
        const int i0 = insert_zero(p, t), i1 = i0 | (1 << t);
        if ((i0 & cl) != cl) continue;
        const double2 s0 = to_d(tile[0][i0]), s1 = to_d(tile[0][i1]);
        const double2 l0 = to_d(tile[1][i0]), l1 = to_d(tile[1][i1]);
        const double2 g0 = cfma(m01, s1, cmul(m00, s0));
        const double2 g1 = cfma(m11, s1, cmul(m10, s0));
        acc += l0.x * g0.x + l0.y * g0.y + l1.x * g1.x + l1.y * g1.y;
    }
    return acc;
}

template <typename T>
__device__ inline double op_grad_diag(typename VT<T>::c **tile, int m, const DevOp &op,
                                      const typename VT<T>::c *Dg) {
    const double2 d0 = to_d(Dg[0]), d1 = to_d(Dg[1]);
    const int t = op.pos[0];
    const int size = 1 << m;
    const int cl = (int)op.cloc;
    double acc = 0;
    for (int i = threadIdx.x; i < size; i += blockDim.x) {
        if ((i & cl) != cl) continue;
        const double2 s = to_d(tile[0][i]), l = to_d(tile[1][i]);
        const double2 g = cmul((i >> t & 1) ? d1 : d0, s);
        acc += l.x * g.x + l.y * g.y;
    }
    return acc;
}

}  // namespace

// ---------------------------------------------------------------------------
// Stage kernel
// ---------------------------------------------------------------------------
// HEAVY = false compiles only the 1-2 qubit, diagonal, scalar and gradient
// micro-ops (low register count, high occupancy for streaming stages);
// HEAVY = true adds the k = 3..5 VALU blocks and the MFMA blocks.
template <typename T, int NV, bool HEAVY>
__global__ void __launch_bounds__(kThreads) k_stage(const StageTask *__restrict__ tasks,
                                                    const int32_t *__restrict__ wg_task, StageTask single) {
    using C = typename VT<T>::c;
    using V = typename VT<T>::v;
    constexpr int PV = VT<T>::per_vec;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double gacc[kMaxGradLDS * (kThreads / 64)];
    __shared__ int perm_lo[256], perm_hi[256];
    __shared__ uint64_t hoff[256];
    __shared__ StageTask tk;

    {
        // copy the task into LDS in 4-byte words, one word per thread
        constexpr int W = (int)(sizeof(StageTask) / 4);
        const uint32_t *src = tasks ? reinterpret_cast<const uint32_t *>(tasks + wg_task[blockIdx.x])
                                    : reinterpret_cast<const uint32_t *>(&single);
        for (int i = threadIdx.x; i < W; i += blockDim.x) reinterpret_cast<uint32_t *>(&tk)[i] = src[i];
    }
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x - tk.wg_begin;
    if (t < 0 || t >= tk.st.n_tiles) return;  // uniform over the workgroup

    const int m = tk.st.m;
    const int size = 1 << m;
    C *tile[2];
    tile[0] = reinterpret_cast<C *>(smem);
    tile[1] = tile[0] + size;
    C *svp[2] = {reinterpret_cast<C *>(tk.sv0), reinterpret_cast<C *>(tk.sv1)};
    const int L = tk.st.L, nh = tk.st.nh;
    const uint64_t lowm = (1ull << L) - 1;
    const uint64_t all = (tk.n >= 64) ? ~0ull : ((1ull << tk.n) - 1);
    const uint64_t free_mask = ~tk.st.tile_mask & all;
    const C *mats = reinterpret_cast<const C *>(tk.mats);
    const DevOp *ops = tk.ops;

    for (int i = threadIdx.x; i < tk.st.n_grad * (kThreads / 64); i += blockDim.x) gacc[i] = 0.0;
    for (int i = threadIdx.x; i < (1 << tk.st.nh) && tk.st.nh <= 8; i += blockDim.x) {
        uint64_t o = 0;
        for (int j = 0; j < tk.st.nh; ++j)
            if (i >> j & 1) o |= 1ull << tk.st.hbits[j];
        hoff[i] = o;
    }
    const bool permuted = tk.st.permuted != 0;
    if (permuted) {
        // local index of store-order index e = perm_lo[e & 255] | perm_hi[e >> 8]
        const int mlo = m < 8 ? m : 8;
        for (int i = threadIdx.x; i < 256; i += blockDim.x) {
            int lo = 0, hi = 0;
            for (int j = 0; j < mlo; ++j)
                if (i >> j & 1) lo |= 1 << tk.st.sperm[j];
            for (int j = 8; j < m; ++j)
                if (i >> (j - 8) & 1) hi |= 1 << tk.st.sperm[j];
            perm_lo[i] = lo;
            perm_hi[i] = hi;
        }
    }

    uint64_t base = 0;
    {
        uint64_t tt = (uint64_t)t, fm = free_mask;
        while (fm) {
            const uint64_t lb = fm & (~fm + 1);
            if (tt & 1) base |= lb;
            tt >>= 1;
            fm ^= lb;
        }
    }
    __syncthreads();  // hoff / perm tables ready
    const bool small_nh = nh <= 8;
    auto seg_off = [&](uint64_t s) -> uint64_t {
        if (small_nh) return hoff[s];
        uint64_t o = 0;
        for (int j = 0; j < nh; ++j)
            if (s >> j & 1) o |= 1ull << tk.st.hbits[j];
        return o;
    };
    // load
    for (int e = threadIdx.x * PV; e < size; e += blockDim.x * PV) {
        const uint64_t s = (uint64_t)e >> L;
        const uint64_t off = base | ((uint64_t)e & lowm) | seg_off(s);
#pragma unroll
        for (int v = 0; v < NV; ++v)
            *reinterpret_cast<V *>(&tile[v][e]) = *reinterpret_cast<const V *>(&svp[v][off]);
    }
    __syncthreads();
    int gi = 0;
    for (int o = tk.st.op_begin; o < tk.st.op_end; ++o) {
        const DevOp op = ops[o];
        const bool is_grad = op.kind == OPK_GRAD || op.kind == OPK_GRADDIAG;
        if ((base & op.cglob) != op.cglob) {
            gi += is_grad;
            continue;
        }
        const C *M = mats + op.mat;
        switch (op.kind) {
            case OPK_DENSE1: op_dense1<T, NV>(tile, m, op, M); break;
            case OPK_DENSE2: op_dense2<T, NV>(tile, m, op, M); break;
            case OPK_DENSEK:
                if constexpr (HEAVY) {
                    if (op.k == 3) op_densek<T, NV, 3>(tile, m, op, M);
                    else if (op.k == 4) op_densek<T, NV, 4>(tile, m, op, M);
                    else op_densek<T, NV, 5>(tile, m, op, M);
                }
                break;
            case OPK_MFMA:
                if constexpr (HEAVY) {
                    if (op.k == 4) op_mfma<T, NV, 4>(tile, m, op, M);
                    else op_mfma<T, NV, 5>(tile, m, op, M);
                }
                break;
            case OPK_DIAG: op_diag<T, NV>(tile, m, op, M, base); break;
            case OPK_SCALAR: op_scalar<T, NV>(tile, m, M); break;
            case OPK_GRAD:
            case OPK_GRADDIAG:
                if constexpr (NV == 2) {
                    const double part = op.kind == OPK_GRAD ? op_grad_dense<T>(tile, m, op, M)
                                                            : op_grad_diag<T>(tile, m, op, M);
                    // one wave-level reduction; each wave owns its own slot
                    const double ws = wave_sum(part);
                    if ((threadIdx.x & 63) == 0) gacc[gi * (kThreads / 64) + (threadIdx.x >> 6)] = ws;
                }
                ++gi;
                break;
            default: break;
        }
        __syncthreads();
    }
    // store (optionally through the write-back bit permutation)
    for (int e = threadIdx.x * PV; e < size; e += blockDim.x * PV) {
        const uint64_t s = (uint64_t)e >> L;
        const uint64_t off = base | ((uint64_t)e & lowm) | seg_off(s);
        if (!permuted) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
                *reinterpret_cast<V *>(&svp[v][off]) = *reinterpret_cast<const V *>(&tile[v][e]);
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                C tmp[PV];
#pragma unroll
                for (int i = 0; i < PV; ++i) {
                    const int ee = e + i;
                    tmp[i] = tile[v][perm_lo[ee & 255] | perm_hi[ee >> 8]];
                }
                *reinterpret_cast<V *>(&svp[v][off]) = *reinterpret_cast<const V *>(tmp);
            }
        }
    }
    if (tk.st.n_grad > 0) {
        __syncthreads();
        int g2 = 0;
        for (int o = tk.st.op_begin; o < tk.st.op_end; ++o) {
            const int kind = ops[o].kind;
            if (kind == OPK_GRAD || kind == OPK_GRADDIAG) {
                if (threadIdx.x == 0) {
                    double sum = 0;
                    for (int w = 0; w < kThreads / 64; ++w) sum += gacc[g2 * (kThreads / 64) + w];
                    tk.partials[(int64_t)ops[o].slot * tk.pstride + t] = sum;
                }
                ++g2;
            }
        }
    }
}

template <typename T>
hipError_t launch_stage(int nv, bool heavy, const StageTask *tasks, const int32_t *wg_task, const StageTask &single,
                        int64_t grid, int max_m, hipStream_t s) {
    using C = typename VT<T>::c;
    const size_t lds = (size_t)nv * ((size_t)1 << max_m) * sizeof(C);
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffffll) return hipErrorInvalidValue;
    const dim3 g((unsigned)grid);
    if (heavy) {
        if (nv == 1) hipLaunchKernelGGL((k_stage<T, 1, true>), g, dim3(kThreads), lds, s, tasks, wg_task, single);
        else hipLaunchKernelGGL((k_stage<T, 2, true>), g, dim3(kThreads), lds, s, tasks, wg_task, single);
    } else {
        if (nv == 1) hipLaunchKernelGGL((k_stage<T, 1, false>), g, dim3(kThreads), lds, s, tasks, wg_task, single);
        else hipLaunchKernelGGL((k_stage<T, 2, false>), g, dim3(kThreads), lds, s, tasks, wg_task, single);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Multi-instance helpers for batches
// ---------------------------------------------------------------------------
template <typename Task>
__device__ inline int find_task(const Task *tasks, int ntasks) {
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].wg_begin <= (int64_t)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_matvec_tasks(const MatvecTask *__restrict__ tasks, int ntasks) {
    using C = typename VT<T>::c;
    __shared__ double red[kThreads / 64];
    const MatvecTask tk = tasks[find_task(tasks, ntasks)];
    const int64_t w = (int64_t)blockIdx.x - tk.wg_begin;
    const int64_t dim = 1ll << tk.n;
    const int64_t i0 = w * tk.chunk;
    const int64_t i1 = i0 + tk.chunk < dim ? i0 + tk.chunk : dim;
    const C *psi = reinterpret_cast<const C *>(tk.psi);
    C *lam = reinterpret_cast<C *>(tk.lam);
    C *psi2 = reinterpret_cast<C *>(tk.psi2);
    const MatvecGroup *groups = reinterpret_cast<const MatvecGroup *>(tk.groups);
    double e = 0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        double2 acc = make_double2(0.0, 0.0);
        for (int g = 0; g < tk.ngroups; ++g) {
            const MatvecGroup gr = groups[g];
            const uint64_t s = (uint64_t)i ^ gr.x;
            double fr = 0, fi = 0;
            for (int q = gr.begin; q < gr.end; ++q) {
                const double2 c = tk.cc[q];
                if (__popcll(s & tk.z[q]) & 1) {
                    fr -= c.x;
                    fi -= c.y;
                } else {
                    fr += c.x;
                    fi += c.y;
                }
            }
            const double2 a = to_d(psi[s]);
            acc.x += fr * a.x - fi * a.y;
            acc.y += fr * a.y + fi * a.x;
        }
        const C p = psi[i];
        const double2 pd = to_d(p);
        e += pd.x * acc.x + pd.y * acc.y;
        C r;
        r.x = (T)acc.x;
        r.y = (T)acc.y;
        lam[i] = r;
        if (psi2) psi2[i] = p;
    }
    const double sum = block_sum(e, red);
    if (threadIdx.x == 0) tk.epart[w] = sum;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_fill_tasks(const FillTask *__restrict__ tasks, int ntasks) {
    using C = typename VT<T>::c;
    const FillTask tk = tasks[find_task(tasks, ntasks)];
    const int64_t w = (int64_t)blockIdx.x - tk.wg_begin;
    const int64_t i0 = w * tk.chunk;
    const int64_t i1 = i0 + tk.chunk < tk.dim ? i0 + tk.chunk : tk.dim;
    C *dst = reinterpret_cast<C *>(tk.dst);
    const C *src = reinterpret_cast<const C *>(tk.src);
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        C v;
        if (src) {
            v = src[i];
        } else {
            v.x = i == 0 ? T(1) : T(0);
            v.y = T(0);
        }
        dst[i] = v;
    }
}

__global__ void __launch_bounds__(kThreads) k_sum_tasks(const SumTask *__restrict__ tasks) {
    __shared__ double red[kThreads / 64];
    const SumTask tk = tasks[blockIdx.x];
    double acc = 0;
    for (int64_t i = threadIdx.x; i < tk.count; i += blockDim.x) acc += tk.src[i];
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) *tk.dst = s;
}

template <typename T>
hipError_t launch_matvec_tasks(const MatvecTask *tasks, int ntasks, int64_t grid, hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_matvec_tasks<T>), dim3((unsigned)grid), dim3(kThreads), 0, s, tasks, ntasks);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_fill_tasks(const FillTask *tasks, int ntasks, int64_t grid, hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_fill_tasks<T>), dim3((unsigned)grid), dim3(kThreads), 0, s, tasks, ntasks);
    return hipGetLastError();
}
hipError_t launch_sum_tasks(const SumTask *tasks, int ntasks, hipStream_t s) {
    if (ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_tasks, dim3(ntasks), dim3(kThreads), 0, s, tasks);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Direct single-op kernel (no LDS): one dense 1-/2-qubit or diagonal op over
// the whole state, streaming.  Thread g owns VEC consecutive groups; when bit
// 0 is neither a target nor a control, VEC = 2 amplitudes of c64 share one
// 16-byte access, so every load/store of a wave is a contiguous run.
// `pos` are global bit positions (matrix index bit j <-> pos[j]).
// ---------------------------------------------------------------------------
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
template <typename V> struct NtVec;
template <> struct NtVec<float4> { using t = nt_f4; };
template <> struct NtVec<float2> { using t = nt_f2; };
template <> struct NtVec<double2> { using t = nt_d2; };

// Plain or non-temporal (streaming) vector access.
template <typename V, bool NT> __device__ inline V ld(const V *p) {
    if constexpr (NT) {
        using N = typename NtVec<V>::t;
        const N v = __builtin_nontemporal_load(reinterpret_cast<const N *>(p));
        V r;
        __builtin_memcpy(&r, &v, sizeof(V));
        return r;
    } else {
        return *p;
    }
}
template <typename V, bool NT> __device__ inline void st(V *p, V v) {
    if constexpr (NT) {
        using N = typename NtVec<V>::t;
        N w;
        __builtin_memcpy(&w, &v, sizeof(V));
        __builtin_nontemporal_store(w, reinterpret_cast<N *>(p));
    } else {
        *p = v;
    }
}

template <typename T, int K, int VEC, bool DIAG, int GPT, bool NT>
__global__ void __launch_bounds__(kThreads) k_direct(typename VT<T>::c *__restrict__ sv, int64_t groups, DirectOp op,
                                                    const typename VT<T>::c *__restrict__ M) {
    using C = typename VT<T>::c;
    constexpr int D = 1 << K;
    int srt[K];
#pragma unroll
    for (int j = 0; j < K; ++j) srt[j] = op.pos[j];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int b = 0; b < K - 1; ++b)
            if (srt[b] > srt[b + 1]) { int t = srt[b]; srt[b] = srt[b + 1]; srt[b + 1] = t; }
    uint64_t off[D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        uint64_t o = 0;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (s >> j & 1) o |= 1ull << op.pos[j];
        off[s] = o;
    }
    C mm[DIAG ? D : D * D];
#pragma unroll
    for (int i = 0; i < (DIAG ? D : D * D); ++i) mm[i] = M[i];
    const uint64_t cm = op.cmask;
    // thread t owns group vectors t, t + T, ..., t + (GPT-1) T (T = threads in
    // the grid): every access instruction of a wave is one contiguous run.
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t b[GPT];
    bool on[GPT];
    C a[GPT][D][VEC];
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
        const int64_t g0 = (t0 + u * nthr) * VEC;
        uint64_t x = (uint64_t)g0;
#pragma unroll
        for (int j = 0; j < K; ++j) x = ((x >> srt[j]) << (srt[j] + 1)) | (x & ((1ull << srt[j]) - 1));
        b[u] = x;
        on[u] = g0 < groups && (x & cm) == cm;
        if (!on[u]) continue;
#pragma unroll
        for (int s = 0; s < D; ++s) {
            if constexpr (VEC == 2) {
                const float4 v = ld<float4, NT>(reinterpret_cast<const float4 *>(&sv[x | off[s]]));
                a[u][s][0] = make_float2(v.x, v.y);
                a[u][s][1] = make_float2(v.z, v.w);
            } else {
                a[u][s][0] = ld<C, NT>(&sv[x | off[s]]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
        if (!on[u]) continue;
        C r[D][VEC];
#pragma unroll
        for (int w = 0; w < VEC; ++w)
#pragma unroll
            for (int q = 0; q < D; ++q) {
                if constexpr (DIAG) {
                    r[q][w] = cmul(mm[q], a[u][q][w]);
                } else {
                    C acc = cmul(mm[q * D], a[u][0][w]);
#pragma unroll
                    for (int c = 1; c < D; ++c) acc = cfma(mm[q * D + c], a[u][c][w], acc);
                    r[q][w] = acc;
                }
            }
#pragma unroll
        for (int s = 0; s < D; ++s) {
            if constexpr (VEC == 2) {
                st<float4, NT>(reinterpret_cast<float4 *>(&sv[b[u] | off[s]]),
                               make_float4(r[s][0].x, r[s][0].y, r[s][1].x, r[s][1].y));
            } else {
                st<C, NT>(&sv[b[u] | off[s]], r[s][0]);
            }
        }
    }
}

namespace {
// Tuning knobs of the direct kernel (group vectors per thread, non-temporal
// access), read once from DESV_DIRECT_GPT / DESV_DIRECT_NT.
int direct_gpt() {
    static int v = [] {
        const char *e = std::getenv("DESV_DIRECT_GPT");
        const int x = e ? std::atoi(e) : 1;
        return (x == 1 || x == 2 || x == 4) ? x : 1;
    }();
    return v;
}
bool direct_nt() {
    static bool v = [] {
        const char *e = std::getenv("DESV_DIRECT_NT");
        return e ? std::atoi(e) != 0 : false;
    }();
    return v;
}
}  // namespace

template <typename T, int K, int VEC, bool DIAG>
hipError_t launch_direct_k(typename VT<T>::c *p, int64_t groups, const DirectOp &op, const typename VT<T>::c *m,
                           hipStream_t s) {
    const int gpt = direct_gpt();
    const bool nt = direct_nt();
    const int64_t vecs = (groups + VEC - 1) / VEC;
    const int64_t threads = (vecs + gpt - 1) / gpt;
    const int64_t grid = std::max<int64_t>(1, (threads + kThreads - 1) / kThreads);
    if (grid > 0x7fffffffll) return hipErrorInvalidValue;
#define DESV_DK(G, N) \
    hipLaunchKernelGGL((k_direct<T, K, VEC, DIAG, G, N>), dim3((unsigned)grid), dim3(kThreads), 0, s, p, groups, op, m)
    if (nt) {
        if (gpt == 1) DESV_DK(1, true); else if (gpt == 2) DESV_DK(2, true); else DESV_DK(4, true);
    } else {
        if (gpt == 1) DESV_DK(1, false); else if (gpt == 2) DESV_DK(2, false); else DESV_DK(4, false);
    }
#undef DESV_DK
    return hipGetLastError();
}

template <typename T>
hipError_t launch_direct(void *sv, int n, const DirectOp &op, const void *M, hipStream_t s) {
    using C = typename VT<T>::c;
    const int K = op.k;
    const int64_t groups = 1ll << (n - K);
    bool bit0 = (op.cmask & 1) != 0;
    for (int j = 0; j < K; ++j) bit0 |= op.pos[j] == 0;
    const bool vec = std::is_same<T, float>::value && !bit0 && groups >= 2;
    C *p = reinterpret_cast<C *>(sv);
    const C *m = reinterpret_cast<const C *>(M);
    if constexpr (std::is_same<T, float>::value) {
        if (K == 1) {
            if (op.diag) return vec ? launch_direct_k<T, 1, 2, true>(p, groups, op, m, s)
                                    : launch_direct_k<T, 1, 1, true>(p, groups, op, m, s);
            return vec ? launch_direct_k<T, 1, 2, false>(p, groups, op, m, s)
                       : launch_direct_k<T, 1, 1, false>(p, groups, op, m, s);
        }
        if (op.diag) return vec ? launch_direct_k<T, 2, 2, true>(p, groups, op, m, s)
                                : launch_direct_k<T, 2, 1, true>(p, groups, op, m, s);
        return vec ? launch_direct_k<T, 2, 2, false>(p, groups, op, m, s)
                   : launch_direct_k<T, 2, 1, false>(p, groups, op, m, s);
    } else {
        if (K == 1) return op.diag ? launch_direct_k<T, 1, 1, true>(p, groups, op, m, s)
                                   : launch_direct_k<T, 1, 1, false>(p, groups, op, m, s);
        return op.diag ? launch_direct_k<T, 2, 1, true>(p, groups, op, m, s)
                       : launch_direct_k<T, 2, 1, false>(p, groups, op, m, s);
    }
}

size_t stage_lds_limit(int device) {
    static size_t cached[64] = {0};
    if (device < 0 || device >= 64) return 64 * 1024;
    if (cached[device]) return cached[device];
    int v = 0;
    size_t lim = 64 * 1024;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
        lim = (size_t)v;
    // leave room for the static gradient/reduction arrays
    lim -= kMaxGradLDS * (kThreads / 64) * sizeof(double) + 512 * sizeof(int) + 256 * 8 + sizeof(StageTask);
    const void *fns[8] = {(const void *)k_stage<float, 1, false>,  (const void *)k_stage<float, 2, false>,
                          (const void *)k_stage<double, 1, false>, (const void *)k_stage<double, 2, false>,
                          (const void *)k_stage<float, 1, true>,   (const void *)k_stage<float, 2, true>,
                          (const void *)k_stage<double, 1, true>,  (const void *)k_stage<double, 2, true>};
    for (const void *f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim) != hipSuccess) {
            (void)hipGetLastError();
            lim = 64 * 1024 - kMaxGradLDS * (kThreads / 64) * sizeof(double) - 512 * sizeof(int) - 256 * 8 - sizeof(StageTask);
        }
    cached[device] = lim;
    return lim;
}

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill_basis(typename VT<T>::c *sv, int64_t dim, int64_t index) {
    using C = typename VT<T>::c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        C z;
        z.x = (i == index) ? T(1) : T(0);
        z.y = T(0);
        sv[i] = z;
    }
}

template <typename T>
__global__ void k_scale(typename VT<T>::c *sv, int64_t dim, double re, double im) {
    using C = typename VT<T>::c;
    C f;
    f.x = (T)re;
    f.y = (T)im;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x)
        sv[i] = cmul(f, sv[i]);
}

inline int ew_grid(int64_t dim) {
    int64_t g = (dim + kThreads - 1) / kThreads;
    if (g > 65536) g = 65536;
    return (int)(g < 1 ? 1 : g);
}

template <typename T>
hipError_t launch_fill_basis(void *sv, int64_t dim, int64_t index, hipStream_t s) {
    hipLaunchKernelGGL((k_fill_basis<T>), dim3(ew_grid(dim)), dim3(kThreads), 0, s, (typename VT<T>::c *)sv, dim, index);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_scale(void *sv, int64_t dim, double re, double im, hipStream_t s) {
    hipLaunchKernelGGL((k_scale<T>), dim3(ew_grid(dim)), dim3(kThreads), 0, s, (typename VT<T>::c *)sv, dim, re, im);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 reductions
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kThreads) k_norm2(const typename VT<T>::c *sv, int64_t dim, double *partials) {
    __shared__ double red[kThreads / 64];
    double acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d(sv[i]);
        acc += a.x * a.x + a.y * a.y;
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_inner(const typename VT<T>::c *a, const typename VT<T>::c *b,
                                                    int64_t dim, double *partials) {
    __shared__ double red[kThreads / 64];
    double re = 0, im = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 x = to_d(a[i]), y = to_d(b[i]);
        re += x.x * y.x + x.y * y.y;
        im += x.x * y.y - x.y * y.x;
    }
    const double sr = block_sum(re, red);
    const double si = block_sum(im, red);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = sr;
        partials[gridDim.x + blockIdx.x] = si;
    }
}

template <typename T>
__global__ void k_probs(const typename VT<T>::c *sv, int64_t dim, double *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d(sv[i]);
        out[i] = a.x * a.x + a.y * a.y;
    }
}

template <typename T>
__global__ void __launch_bounds__(kThreads) k_qubit_probs(const typename VT<T>::c *sv, int64_t dim, int bit,
                                                          double *partials) {
    __shared__ double red[kThreads / 64];
    double p0 = 0, p1 = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 a = to_d(sv[i]);
        const double p = a.x * a.x + a.y * a.y;
        if ((i >> bit) & 1) p1 += p;
        else p0 += p;
    }
    const double s0 = block_sum(p0, red);
    const double s1 = block_sum(p1, red);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s0;
        partials[gridDim.x + blockIdx.x] = s1;
    }
}

template <typename T>
__global__ void k_collapse(typename VT<T>::c *sv, int64_t dim, int bit, int outcome, double scale) {
    using C = typename VT<T>::c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        C a = sv[i];
        if ((int)((i >> bit) & 1) != outcome) {
            a.x = T(0);
            a.y = T(0);
        } else {
            a.x = (T)((double)a.x * scale);
            a.y = (T)((double)a.y * scale);
        }
        sv[i] = a;
    }
}

__global__ void __launch_bounds__(kThreads) k_sum_partials(const double *partials, int nblk, double *out) {
    __shared__ double red[kThreads / 64];
    const double *p = partials + (int64_t)blockIdx.x * nblk;
    double acc = 0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) acc += p[i];
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

template <typename T>
hipError_t launch_norm2(const void *sv, int64_t dim, double *partials, int nblk, hipStream_t s) {
    hipLaunchKernelGGL((k_norm2<T>), dim3(nblk), dim3(kThreads), 0, s, (const typename VT<T>::c *)sv, dim, partials);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_inner(const void *a, const void *b, int64_t dim, double *partials, int nblk, hipStream_t s) {
    hipLaunchKernelGGL((k_inner<T>), dim3(nblk), dim3(kThreads), 0, s, (const typename VT<T>::c *)a,
                       (const typename VT<T>::c *)b, dim, partials);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_probs(const void *sv, int64_t dim, double *out, hipStream_t s) {
    hipLaunchKernelGGL((k_probs<T>), dim3(ew_grid(dim)), dim3(kThreads), 0, s, (const typename VT<T>::c *)sv, dim, out);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_qubit_probs(const void *sv, int64_t dim, int bit, double *partials, int nblk, hipStream_t s) {
    hipLaunchKernelGGL((k_qubit_probs<T>), dim3(nblk), dim3(kThreads), 0, s, (const typename VT<T>::c *)sv, dim, bit,
                       partials);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_collapse(void *sv, int64_t dim, int bit, int outcome, double scale, hipStream_t s) {
    hipLaunchKernelGGL((k_collapse<T>), dim3(ew_grid(dim)), dim3(kThreads), 0, s, (typename VT<T>::c *)sv, dim, bit,
                       outcome, scale);
    return hipGetLastError();
}
hipError_t launch_sum_partials(const double *partials, int nvals, int nblk, double *out, hipStream_t s) {
    if (nvals <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_partials, dim3(nvals), dim3(kThreads), 0, s, partials, nblk, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pauli sums (reference physics/observables.py:63-149, :250-308)
// ---------------------------------------------------------------------------
// <psi|P|psi> for all terms of one flip mask x.  For x != 0 the state is
// walked once per pair (i, j = i^x) with bit h (top bit of x) clear in i:
//   conj(a_i) s_j a_j + conj(a_j) s_i a_i = (s_i + s_j) Re w + i (s_j - s_i) Im w,
// w = conj(a_i) a_j, s = (-1)^popcount(index & z).  The i^{#Y} factor is
// applied on the host.  Partials: [2*term + {re,im}][nblk].
template <typename T>
__global__ void __launch_bounds__(kThreads) k_pauli_expect(const typename VT<T>::c *sv, int n, const PauliItem *items,
                                                           const uint64_t *zmask, double *partials, int nblk) {
    __shared__ double red[kThreads / 64];
    const PauliItem it = items[blockIdx.y];
    const uint64_t x = it.x;
    const int nt = it.count;
    uint64_t z[kTermsPerItem];
#pragma unroll
    for (int t = 0; t < kTermsPerItem; ++t) z[t] = t < nt ? zmask[it.begin + t] : 0;
    double ar[kTermsPerItem], ai[kTermsPerItem];
#pragma unroll
    for (int t = 0; t < kTermsPerItem; ++t) ar[t] = ai[t] = 0;
    const int64_t dim = 1ll << n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (x == 0) {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += stride) {
            const double2 a = to_d(sv[i]);
            const double p = a.x * a.x + a.y * a.y;
#pragma unroll
            for (int t = 0; t < kTermsPerItem; ++t)
                ar[t] += (__popcll((uint64_t)i & z[t]) & 1) ? -p : p;
        }
    } else {
        const int h = 63 - __clzll(x);
        const int64_t half = dim >> 1;
        for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < half; q += stride) {
            const uint64_t i = (((uint64_t)q >> h) << (h + 1)) | ((uint64_t)q & ((1ull << h) - 1));
            const uint64_t j = i ^ x;
            const double2 ai_ = to_d(sv[i]), aj = to_d(sv[j]);
            const double wr = ai_.x * aj.x + ai_.y * aj.y;
            const double wi = ai_.x * aj.y - ai_.y * aj.x;
#pragma unroll
            for (int t = 0; t < kTermsPerItem; ++t) {
                const double si = (__popcll(i & z[t]) & 1) ? -1.0 : 1.0;
                const double sj = (__popcll(j & z[t]) & 1) ? -1.0 : 1.0;
                ar[t] += (si + sj) * wr;
                ai[t] += (sj - si) * wi;
            }
        }
    }
    for (int t = 0; t < nt; ++t) {
        const double sr = block_sum(ar[t], red);
        const double si = block_sum(ai[t], red);
        if (threadIdx.x == 0) {
            partials[(int64_t)(2 * (it.begin + t)) * nblk + blockIdx.x] = sr;
            partials[(int64_t)(2 * (it.begin + t) + 1) * nblk + blockIdx.x] = si;
        }
    }
}

// out_i = sum_groups (sum_{t in g} cc_t (-1)^popcount((i^x)&z_t)) in_{i^x}, FP64 accumulation.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_pauli_matvec(typename VT<T>::c *out, const typename VT<T>::c *in, int n,
                                                           const MatvecGroup *groups, int n_groups,
                                                           const uint64_t *zmask, const double2 *cc) {
    using C = typename VT<T>::c;
    const int64_t dim = 1ll << n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = make_double2(0.0, 0.0);
        for (int g = 0; g < n_groups; ++g) {
            const MatvecGroup gr = groups[g];
            const uint64_t s = (uint64_t)i ^ gr.x;
            double fr = 0, fi = 0;
            for (int t = gr.begin; t < gr.end; ++t) {
                const double2 c = cc[t];
                if (__popcll(s & zmask[t]) & 1) {
                    fr -= c.x;
                    fi -= c.y;
                } else {
                    fr += c.x;
                    fi += c.y;
                }
            }
            const double2 a = to_d(in[s]);
            acc.x += fr * a.x - fi * a.y;
            acc.y += fr * a.y + fi * a.x;
        }
        C r;
        r.x = (T)acc.x;
        r.y = (T)acc.y;
        out[i] = r;
    }
}

template <typename T>
__global__ void k_dense_matvec(typename VT<T>::c *out, const typename VT<T>::c *in, int n, const double2 *H) {
    using C = typename VT<T>::c;
    const int64_t dim = 1ll << n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = make_double2(0.0, 0.0);
        const double2 *row = H + i * dim;
        for (int64_t j = 0; j < dim; ++j) acc = cfma(row[j], to_d(in[j]), acc);
        C r;
        r.x = (T)acc.x;
        r.y = (T)acc.y;
        out[i] = r;
    }
}

template <typename T>
hipError_t launch_pauli_expect(const void *sv, int n, const PauliItem *items, int n_items, const uint64_t *zmask,
                               double *partials, int nblk, hipStream_t s) {
    hipLaunchKernelGGL((k_pauli_expect<T>), dim3(nblk, n_items), dim3(kThreads), 0, s,
                       (const typename VT<T>::c *)sv, n, items, zmask, partials, nblk);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_pauli_matvec(void *out, const void *in, int n, const MatvecGroup *groups, int n_groups,
                               const uint64_t *zmask, const double2 *cc, hipStream_t s) {
    hipLaunchKernelGGL((k_pauli_matvec<T>), dim3(ew_grid(1ll << n)), dim3(kThreads), 0, s,
                       (typename VT<T>::c *)out, (const typename VT<T>::c *)in, n, groups, n_groups, zmask, cc);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_dense_matvec(void *out, const void *in, int n, const double2 *H, hipStream_t s) {
    hipLaunchKernelGGL((k_dense_matvec<T>), dim3(ew_grid(1ll << n)), dim3(kThreads), 0, s,
                       (typename VT<T>::c *)out, (const typename VT<T>::c *)in, n, H);
    return hipGetLastError();
}

#define DESV_INST(T)                                                                                            \
    template hipError_t launch_direct<T>(void *, int, const DirectOp &, const void *, hipStream_t);            \
    template hipError_t launch_stage<T>(int, bool, const StageTask *, const int32_t *, const StageTask &, int64_t, int, \
                                        hipStream_t);                                                          \
    template hipError_t launch_matvec_tasks<T>(const MatvecTask *, int, int64_t, hipStream_t);                 \
    template hipError_t launch_fill_tasks<T>(const FillTask *, int, int64_t, hipStream_t);                     \
    template hipError_t launch_fill_basis<T>(void *, int64_t, int64_t, hipStream_t);                          \
    template hipError_t launch_scale<T>(void *, int64_t, double, double, hipStream_t);                        \
    template hipError_t launch_norm2<T>(const void *, int64_t, double *, int, hipStream_t);                   \
    template hipError_t launch_inner<T>(const void *, const void *, int64_t, double *, int, hipStream_t);     \
    template hipError_t launch_probs<T>(const void *, int64_t, double *, hipStream_t);                        \
    template hipError_t launch_qubit_probs<T>(const void *, int64_t, int, double *, int, hipStream_t);        \
    template hipError_t launch_collapse<T>(void *, int64_t, int, int, double, hipStream_t);                   \
    template hipError_t launch_pauli_expect<T>(const void *, int, const PauliItem *, int, const uint64_t *,   \
                                               double *, int, hipStream_t);                                   \
    template hipError_t launch_pauli_matvec<T>(void *, const void *, int, const MatvecGroup *, int,           \
                                               const uint64_t *, const double2 *, hipStream_t);               \
    template hipError_t launch_dense_matvec<T>(void *, const void *, int, const double2 *, hipStream_t);

DESV_INST(float)
DESV_INST(double)

}  // namespace desv
