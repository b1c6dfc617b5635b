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

// Complex products written as explicit FMA chains: a complex multiply-add is
// four FMAs (the naive `acc += a*b - c*d` costs a multiply, an FMA and an add
// per component).
__device__ inline float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename C> __device__ inline C cmul(C a, C b) {
    C r;
    r.x = fmadd(-a.y, b.y, a.x * b.x);
    r.y = fmadd(a.y, b.x, a.x * b.y);
    return r;
}
template <typename C> __device__ inline C cfma(C a, C b, C acc) {
    acc.x = fmadd(-a.y, b.y, fmadd(a.x, b.x, acc.x));
    acc.y = fmadd(a.y, b.x, fmadd(a.x, b.y, acc.y));
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
// LDS tile layout.  Tile-local amplitude i lives at element slot(i) =
// i ^ ((i >> 5) & 31): index bits 5..9 are folded onto bits 0..4, so a wave
// whose lanes step through the tile with any stride 2^s, s <= 5, touches
// distinct banks.  slot() is linear over XOR (slot(a ^ b) = slot(a) ^ slot(b))
// and an involution.  Every address below is built from disjoint bit fields,
// so its slot is the XOR of the fields' slots: an op computes one per-lane
// slot and combines it with wave-uniform slots (one VALU XOR per access).
// Offsets are in bytes (slot * sizeof(amplitude)), which is also linear.
// ---------------------------------------------------------------------------
__device__ inline int slot(int i) { return i ^ ((i >> 5) & 31); }

// Scalar (wave-uniform) copies of values that every lane holds equal.
__device__ inline int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline uint64_t uni64(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}
template <typename P> __device__ inline P *uni_ptr(P *p) {
    return reinterpret_cast<P *>(uni64(reinterpret_cast<uint64_t>(p)));
}
// Global-address-space view of a wave-uniform pointer: loads through it are
// global (not flat) loads, and scalar loads when the address is uniform.
#define DESV_GLOBAL __attribute__((address_space(1)))
template <typename P> __device__ inline const DESV_GLOBAL P *gptr(const P *p) {
    return (const DESV_GLOBAL P *)uni_ptr(p);
}
template <typename P> __device__ inline DESV_GLOBAL P *gptr_rw(P *p) { return (DESV_GLOBAL P *)uni_ptr(p); }
// Loads/stores of HIP vector types through global pointers, as plain vectors.
template <typename X> struct RawOf;
template <> struct RawOf<float2> { using t = float __attribute__((ext_vector_type(2))); };
template <> struct RawOf<float4> { using t = float __attribute__((ext_vector_type(4))); };
template <> struct RawOf<double2> { using t = double __attribute__((ext_vector_type(2))); };
template <> struct RawOf<int4> { using t = int __attribute__((ext_vector_type(4))); };
template <typename X> __device__ inline X gload(const DESV_GLOBAL X *p) {
    using R = typename RawOf<X>::t;
    const R r = *(const DESV_GLOBAL R *)p;
    X x;
    __builtin_memcpy(&x, &r, sizeof(X));
    return x;
}
template <typename X> __device__ inline void gstore(DESV_GLOBAL X *p, const X &x) {
    using R = typename RawOf<X>::t;
    R r;
    __builtin_memcpy(&r, &x, sizeof(X));
    *(DESV_GLOBAL R *)p = r;
}
// Constant-address-space view of a wave-uniform pointer to data no kernel of
// the launch writes (op descriptors): every load through it is a scalar load,
// whatever the compiler can prove about aliasing stores.
#define DESV_CONST __attribute__((address_space(4)))
template <typename P> __device__ inline const DESV_CONST P *cptr(const P *p) {
    return (const DESV_CONST P *)uni_ptr(p);
}
__device__ inline DevOp gload_op(const DESV_CONST DevOp *p) {
    constexpr int W = (int)(sizeof(DevOp) / 4);
    const DESV_CONST uint32_t *w = (const DESV_CONST uint32_t *)p;
    uint32_t buf[W];
#pragma unroll
    for (int i = 0; i < W; ++i) buf[i] = w[i];
    DevOp op;
    __builtin_memcpy(&op, buf, sizeof(DevOp));
    return op;
}

// Deposit the bits of x, lowest first, into the bit positions not in `fixed`.
__device__ inline int spread(int x, int fixed) {
    for (int f = fixed; f; f &= f - 1) x = insert_zero(x, __builtin_ctz(f));
    return x;
}

// The stage tile lives at LDS address 0 (all of the stage kernel's LDS is
// dynamic, the bookkeeping behind the tiles): tile pointers are plain
// address-space-3 pointers built from the integer 0, so a slot is the address.
using lds_t = __attribute__((address_space(3))) char;
template <typename C> using lds_of = __attribute__((address_space(3))) C;
template <typename C> __device__ inline C lds_ld(const lds_t *p, int off) {
    using R = typename RawOf<C>::t;
    const R r = *reinterpret_cast<const lds_of<R> *>(p + off);
    C x;
    __builtin_memcpy(&x, &r, sizeof(C));
    return x;
}
template <typename C> __device__ inline void lds_st(lds_t *p, int off, C v) {
    using R = typename RawOf<C>::t;
    R r;
    __builtin_memcpy(&r, &v, sizeof(C));
    *reinterpret_cast<lds_of<R> *>(p + off) = r;
}

// Work split of an op whose group enumerates the tile bits not in `fixed`:
// lane `tid` owns groups tid + 256 j, j < J; lanes with tid >= G idle.
struct Groups {
    int fixed;  // target and control bits
    int G;      // number of groups
    int J;      // groups per lane
    int P;      // byte slot of this lane's first group
    __device__ Groups(int m, int fixed_, int eb) : fixed(fixed_) {
        G = 1 << (m - __builtin_popcount(fixed_));
        J = G > kThreads ? G / kThreads : 1;
        P = slot(spread((int)threadIdx.x, fixed_)) * eb;
    }
    __device__ bool idle() const { return (int)threadIdx.x >= G; }
    // byte slot of the uniform part of group tid + 256 j
    __device__ int U(int j, int eb) const { return slot(spread(j * kThreads, fixed)) * eb; }
};

// ---------------------------------------------------------------------------
// Tile micro-ops.  `lds` = tile 0; tile v is at lds + v * vstride.
// ---------------------------------------------------------------------------

// Compensated ops (CP = true, complex64 programs): the exact matrix is the
// FP32 pair M + ML (ML = the FP32 rounding of the FP64 entries minus M).  Each
// output is accumulated lo part first, so the ML products (about an ulp of the
// result) enter the accumulator before the M products are added; every later
// FMA rounds the exact sum of an unrounded product and the accumulator, so the
// result is M + ML applied with unbiased rounding.  Adding the ML products
// after the M products (or their sum to the rounded M result) would round them
// away, and the rounding error of M would then repeat every time the same
// matrix is applied (a norm drift linear in the number of applications).

// Dense 2^K x 2^K matrix, K = 1, 2, on tile bits pos[] with local controls.
template <typename T, int NV, int K, bool CP, typename MP>
__device__ inline void op_dense_lo(lds_t *lds, int vstride, int m, const DevOp &op, MP M, MP ML) {
    using C = typename VT<T>::c;
    constexpr int D = 1 << K;
    constexpr int EB = sizeof(C);
    const int cl = (int)op.cloc;
    int fixed = cl, so[D];
#pragma unroll
    for (int j = 0; j < K; ++j) fixed |= 1 << op.pos[j];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        int o = cl;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (s >> j & 1) o |= 1 << op.pos[j];
        so[s] = slot(o) * EB;
    }
    const Groups g(m, fixed, EB);
    if (g.idle()) return;
    C mm[D * D], ml[CP ? D * D : 1];
#pragma unroll
    for (int i = 0; i < D * D; ++i) mm[i] = gload(M + i);
    if constexpr (CP) {
#pragma unroll
        for (int i = 0; i < D * D; ++i) ml[i] = gload(ML + i);
    }
    auto row = [&](const C *a, int r) {
        C acc;
        if constexpr (CP) {
            acc = cmul(ml[r * D], a[0]);
#pragma unroll
            for (int c = 1; c < D; ++c) acc = cfma(ml[r * D + c], a[c], acc);
#pragma unroll
            for (int c = 0; c < D; ++c) acc = cfma(mm[r * D + c], a[c], acc);
        } else {
            acc = cmul(mm[r * D], a[0]);
#pragma unroll
            for (int c = 1; c < D; ++c) acc = cfma(mm[r * D + c], a[c], acc);
        }
        return acc;
    };
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        lds_t *tb = lds + v * vstride;
        int j = 0;
        for (; j + 2 <= g.J; j += 2) {
            const int A0 = g.P ^ g.U(j, EB), A1 = g.P ^ g.U(j + 1, EB);
            C a0[D], a1[D];
#pragma unroll
            for (int s = 0; s < D; ++s) {
                a0[s] = lds_ld<C>(tb, A0 ^ so[s]);
                a1[s] = lds_ld<C>(tb, A1 ^ so[s]);
            }
#pragma unroll
            for (int r = 0; r < D; ++r) {
                lds_st<C>(tb, A0 ^ so[r], row(a0, r));
                lds_st<C>(tb, A1 ^ so[r], row(a1, r));
            }
        }
        if (j < g.J) {
            const int A0 = g.P ^ g.U(j, EB);
            C a0[D];
#pragma unroll
            for (int s = 0; s < D; ++s) a0[s] = lds_ld<C>(tb, A0 ^ so[s]);
#pragma unroll
            for (int r = 0; r < D; ++r) lds_st<C>(tb, A0 ^ so[r], row(a0, r));
        }
    }
}

// Dense 2^K x 2^K block on VALU (K = 3..5): one group's amplitudes in
// registers, rows formed one at a time from the (scalar) matrix row.
template <typename T, int NV, int K, bool CP, typename MP>
__device__ inline void op_densek(lds_t *lds, int vstride, int m, const DevOp &op, MP M, MP ML) {
    using C = typename VT<T>::c;
    constexpr int D = 1 << K;
    constexpr int EB = sizeof(C);
    const int cl = (int)op.cloc;
    int fixed = cl;
#pragma unroll
    for (int j = 0; j < K; ++j) fixed |= 1 << op.pos[j];
    // row/column r of the matrix <-> tile offset cl | sum_j bit_j(r) << pos[j]
    auto roff = [&](int r) {
        int o = cl;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (r >> j & 1) o |= 1 << op.pos[j];
        return slot(o) * EB;
    };
    const Groups g(m, fixed, EB);
    if (g.idle()) return;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        lds_t *tb = lds + v * vstride;
        for (int j = 0; j < g.J; ++j) {
            const int A = g.P ^ g.U(j, EB);
            C a[D];
#pragma unroll
            for (int s = 0; s < D; ++s) a[s] = lds_ld<C>(tb, A ^ roff(s));
#pragma unroll 1
            for (int r = 0; r < D; ++r) {
                const MP Mr = M + r * D;
                C acc;
                if constexpr (CP) {
                    const MP Lr = ML + r * D;
                    acc = cmul(gload(Lr), a[0]);
#pragma unroll
                    for (int c = 1; c < D; ++c) acc = cfma(gload(Lr + c), a[c], acc);
#pragma unroll
                    for (int c = 0; c < D; ++c) acc = cfma(gload(Mr + c), a[c], acc);
                } else {
                    acc = cmul(gload(Mr), a[0]);
#pragma unroll
                    for (int c = 1; c < D; ++c) acc = cfma(gload(Mr + c), a[c], acc);
                }
                lds_st<C>(tb, A ^ roff(r), acc);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused k-qubit block (k = 4, 5) as a real-split complex GEMM on MFMA
// (complex64 only).
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

constexpr int kWaves = kThreads / 64;
constexpr int kLogWaves = 2;
static_assert((1 << kLogWaves) == kWaves, "kLogWaves");
// MFMA ops hand column block w + kWaves g (16 groups each) to wave w.  The
// wave-uniform byte slot of block g is looked up rather than recomputed: lane
// l of `lo` holds the slot of g = l (including the wave's own bits), lane l of
// `hi` the slot of g = 64 l, and slot(g) = lo[g % 64] ^ hi[g / 64] (slot and
// spread are linear over XOR on disjoint bit fields).  Lanes past the op's
// block count hold unused values.
struct BlockSlots {
    int lo, hi;
    __device__ BlockSlots(int fixed, int eb) {
        const int lane = threadIdx.x & 63;
        const int wave = uni((int)threadIdx.x >> 6);
        lo = slot(spread((wave << 4) | (lane << (4 + kLogWaves)), fixed)) * eb;
        hi = slot(spread(lane << (10 + kLogWaves), fixed)) * eb;
    }
    struct Given {};
    __device__ BlockSlots(Given, int lo_, int hi_) : lo(lo_), hi(hi_) {}
    __device__ int operator()(int g) const {
        const int l = __builtin_amdgcn_readlane(lo, g & 63);
        return g < 64 ? l : l ^ __builtin_amdgcn_readlane(hi, g >> 6);
    }
};

template <int NV, int K, bool CP, typename MP>
__device__ inline void op_mfma(lds_t *lds, int vstride, int m, const DevOp &op, MP M, MP ML) {
    constexpr int D = 1 << K;
    constexpr int RB = D / 16;  // 16-row blocks of U
    constexpr int KS = D / 4;   // 4-wide k slices
    constexpr int EB = sizeof(float2);
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, k4 = lane >> 4;
    const int cl = (int)op.cloc;
    int fixed = cl;
#pragma unroll
    for (int j = 0; j < K; ++j) fixed |= 1 << op.pos[j];
    auto roff = [&](int r) {
        int o = 0;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (r >> j & 1) o |= 1 << op.pos[j];
        return o;
    };
    // A fragments (constant over the tile)
    float ar[RB][KS], ai[RB][KS], nai[RB][KS];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const float2 u = gload(M + (rb * 16 + i16) * D + ks * 4 + k4);
            ar[rb][ks] = u.x;
            ai[rb][ks] = u.y;
            nai[rb][ks] = -u.y;  // Yr = Ur Xr + (-Ui) Xi: the sign lives on the constant operand
        }
    // lo parts of the compensated form (accumulated first)
    float lr[CP ? RB : 1][CP ? KS : 1], li[CP ? RB : 1][CP ? KS : 1], nli[CP ? RB : 1][CP ? KS : 1];
    if constexpr (CP) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const float2 u = gload(ML + (rb * 16 + i16) * D + ks * 4 + k4);
                lr[rb][ks] = u.x;
                li[rb][ks] = u.y;
                nli[rb][ks] = -u.y;
            }
    }
    // per-lane byte slots: this lane's column (within a 16-column block) and
    // the rows it reads (B) and writes (accumulator)
    const int pl = spread(i16, fixed) | cl;
    int pb[KS], py[RB][4];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) pb[ks] = slot(pl | roff(ks * 4 + k4)) * EB;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int j = 0; j < 4; ++j) py[rb][j] = slot(pl | roff(rb * 16 + 4 * k4 + j)) * EB;

    const int G = 1 << (m - __builtin_popcount(fixed));
    if (NV == 1 && G >= 16 * kWaves) {
        // full column blocks for every wave: no lane predicates, block slots
        // from the lane table (at most 16 blocks a wave, within its low
        // half), software-pipelined like the 3-qubit loop: block it's chains
        // are issued before block it - 1's results are stored and block
        // it + 1's inputs loaded (one vector: with two the prefetched
        // fragments exceed the register budget of four waves per SIMD; plain
        // 5-qubit blocks store each block before the next one's loads, as the
        // held results would not fit beside their 48 fragment registers)
        constexpr bool PIPE = !(K == 5 && !CP);
        const int nit = G / (16 * kWaves);
        const BlockSlots bs(fixed, EB);
        int Uc = __builtin_amdgcn_readlane(bs.lo, 0), Up = Uc;
        float br[NV][KS], bi[NV][KS];
        f32x4_t pr[NV][RB], pi[NV][RB];  // the previous block's results
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const float2 x = lds_ld<float2>(lds + v * vstride, pb[ks] ^ Uc);
                br[v][ks] = x.x;
                bi[v][ks] = x.y;
            }
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) pr[v][rb] = pi[v][rb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        }
        for (int it = 0; it < nit; ++it) {
            f32x4_t yr[NV][RB], yi[NV][RB];
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) {
                    yr[v][rb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
                    yi[v][rb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
                    if constexpr (CP) {
#pragma unroll
                        for (int ks = 0; ks < KS; ++ks) {
                            yr[v][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(lr[rb][ks], br[v][ks], yr[v][rb], 0, 0, 0);
                            yi[v][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(lr[rb][ks], bi[v][ks], yi[v][rb], 0, 0, 0);
                            yr[v][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(nli[rb][ks], bi[v][ks], yr[v][rb], 0, 0, 0);
                            yi[v][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(li[rb][ks], br[v][ks], yi[v][rb], 0, 0, 0);
                        }
                    }
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks) {
                        yr[v][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[rb][ks], br[v][ks], yr[v][rb], 0, 0, 0);
                        yi[v][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[rb][ks], bi[v][ks], yi[v][rb], 0, 0, 0);
                        yr[v][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai[rb][ks], bi[v][ks], yr[v][rb], 0, 0, 0);
                        yi[v][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[rb][ks], br[v][ks], yi[v][rb], 0, 0, 0);
                    }
                }
            if (PIPE && it > 0) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            lds_st<float2>(lds + v * vstride, py[rb][j] ^ Up, make_float2(pr[v][rb][j], pi[v][rb][j]));
            }
            if (!PIPE) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            lds_st<float2>(lds + v * vstride, py[rb][j] ^ Uc, make_float2(yr[v][rb][j], yi[v][rb][j]));
            }
            const int Un = __builtin_amdgcn_readlane(bs.lo, (it + 1) & 63);
            if (it + 1 < nit) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks) {
                        const float2 x = lds_ld<float2>(lds + v * vstride, pb[ks] ^ Un);
                        br[v][ks] = x.x;
                        bi[v][ks] = x.y;
                    }
            }
            if (PIPE) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb) {
                        pr[v][rb] = yr[v][rb];
                        pi[v][rb] = yi[v][rb];
                    }
            }
            Up = Uc;
            Uc = Un;
        }
        if (PIPE) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        lds_st<float2>(lds + v * vstride, py[rb][j] ^ Up, make_float2(pr[v][rb][j], pi[v][rb][j]));
        }
        return;
    }
    const bool valid = i16 < G;  // tiles narrower than 16 columns
    const int step = (blockDim.x >> 6) * 16;
    // UN column blocks per iteration (independent accumulator chains with
    // their LDS loads in flight together); 2 measured no faster than 1
    constexpr int UN = 1;
    for (int c0 = (threadIdx.x >> 6) * 16; c0 < G; c0 += UN * step) {
        int U[UN];
        bool on[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            on[u] = c0 + u * step < G;  // uniform over the wave
            U[u] = on[u] ? slot(spread(c0 + u * step, fixed)) * EB : 0;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            lds_t *tb = lds + v * vstride;
            float br[UN][KS], bi[UN][KS];
#pragma unroll
            for (int u = 0; u < UN; ++u)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const float2 x =
                        (valid && on[u]) ? lds_ld<float2>(tb, pb[ks] ^ U[u]) : make_float2(0.f, 0.f);
                    br[u][ks] = x.x;
                    bi[u][ks] = x.y;
                }
            f32x4_t yr[UN][RB], yi[UN][RB];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
                for (int u = 0; u < UN; ++u) {
                    yr[u][rb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
                    yi[u][rb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
                }
                if constexpr (CP) {
#pragma unroll
                    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                        for (int u = 0; u < UN; ++u) {
                            yr[u][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(lr[rb][ks], br[u][ks], yr[u][rb], 0, 0, 0);
                            yi[u][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(lr[rb][ks], bi[u][ks], yi[u][rb], 0, 0, 0);
                            yr[u][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(nli[rb][ks], bi[u][ks], yr[u][rb], 0, 0, 0);
                            yi[u][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(li[rb][ks], br[u][ks], yi[u][rb], 0, 0, 0);
                        }
                }
#pragma unroll
                for (int ks = 0; ks < KS; ++ks)
#pragma unroll
                    for (int u = 0; u < UN; ++u) {
                        yr[u][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[rb][ks], br[u][ks], yr[u][rb], 0, 0, 0);
                        yi[u][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[rb][ks], bi[u][ks], yi[u][rb], 0, 0, 0);
                        yr[u][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(nai[rb][ks], bi[u][ks], yr[u][rb], 0, 0, 0);
                        yi[u][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[rb][ks], br[u][ks], yi[u][rb], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int u = 0; u < UN; ++u)
                if (valid && on[u]) {
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            lds_st<float2>(tb, py[rb][j] ^ U[u], make_float2(yr[u][rb][j], yi[u][rb][j]));
                }
        }
    }
}

// Fused 3-qubit block on MFMA in the stacked real form: the 16x16 real
// matrix R whose rows interleave the real and imaginary part of each output,
//   R[2k][c] = Ur[k][c],  R[2k][8+c] = -Ui[k][c],
//   R[2k+1][c] = Ui[k][c], R[2k+1][8+c] = Ur[k][c],
// times the real 16-vector [Xr; Xi] of each column: four 16x16x4 MFMAs per
// 16 columns (the same FLOPs as the complex product).  B fragment of step ks:
// lane l -> [Xr; Xi][4 ks + l/16], i.e. rows l/16 and 4 + l/16 of X, real
// parts for ks < 2 and imaginary parts for ks >= 2; accumulator lane l, reg j
// -> R row 4 (l/16) + j, so lane group g holds outputs 2g and 2g+1 as
// (re, im) register pairs.
// This lane's words of a 3-qubit MFMA op's host-built lane table (slots,
// block slots, hi and lo A fragments), or none (`on` false: the op derives
// them in the kernel).  The stage kernel requests an op's words while the
// op before it runs, so their latency hides under that op.
struct LaneWords {
    int4 w0, w1, w2, w3;
    bool on;
};
template <bool CP> __device__ inline LaneWords load_lane_words(const DESV_GLOBAL int4 *lt) {
    LaneWords r;
    const DESV_GLOBAL int4 *w = lt + threadIdx.x * (kLaneTableWords / 4);
    r.w0 = gload(w), r.w1 = gload(w + 1), r.w2 = gload(w + 2);
    r.w3 = CP ? gload(w + 3) : make_int4(0, 0, 0, 0);
    r.on = true;
    return r;
}

template <int NV, bool CP, typename MP>
__device__ inline void op_mfma3(lds_t *lds, int vstride, int m, const DevOp &op, MP M, MP ML, const LaneWords &lw) {
    constexpr int EB = sizeof(float2);
    const int lane = threadIdx.x & 63;
    const int i16 = lane & 15, g4 = lane >> 4;
    const int cl = (int)op.cloc;
    const int p0 = op.pos[0], p1 = op.pos[1], p2 = op.pos[2];
    const int fixed = cl | (1 << p0) | (1 << p1) | (1 << p2);
    auto roff = [&](int r) { return ((r & 1) << p0) | ((r >> 1 & 1) << p1) | ((r >> 2 & 1) << p2); };
    float a[4], la[4];  // A fragments; lo parts of the compensated form (accumulated first)
    int pb0, pb1, py0, py1, blo = 0, bhi = 0;
    if (lw.on) {
        // slots and fragments from the op's host-built lane table
        pb0 = lw.w0.x, pb1 = lw.w0.y, py0 = lw.w0.z, py1 = lw.w0.w, blo = lw.w1.x, bhi = lw.w1.y;
        a[0] = __int_as_float(lw.w2.x), a[1] = __int_as_float(lw.w2.y), a[2] = __int_as_float(lw.w2.z),
        a[3] = __int_as_float(lw.w2.w);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) la[ks] = 0.f;
        if constexpr (CP) {
            la[0] = __int_as_float(lw.w3.x), la[1] = __int_as_float(lw.w3.y), la[2] = __int_as_float(lw.w3.z),
            la[3] = __int_as_float(lw.w3.w);
        }
    } else {
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int kk = 4 * ks + g4;
            const float2 u = gload(M + (i16 >> 1) * 8 + (kk & 7));
            a[ks] = kk < 8 ? ((i16 & 1) ? u.y : u.x) : ((i16 & 1) ? u.x : -u.y);
        }
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            la[ks] = 0.f;
            if constexpr (CP) {
                const int kk = 4 * ks + g4;
                const float2 u = gload(ML + (i16 >> 1) * 8 + (kk & 7));
                la[ks] = kk < 8 ? ((i16 & 1) ? u.y : u.x) : ((i16 & 1) ? u.x : -u.y);
            }
        }
        const int pl = spread(i16, fixed) | cl;
        pb0 = slot(pl | roff(g4)) * EB, pb1 = slot(pl | roff(4 + g4)) * EB;
        py0 = slot(pl | roff(2 * g4)) * EB, py1 = slot(pl | roff(2 * g4 + 1)) * EB;
    }
    // y = R [Xr; Xi] (+ the lo part first when compensated) for one column block
    auto mv = [&](float2 x0, float2 x1) {
        f32x4_t y = f32x4_t{0.f, 0.f, 0.f, 0.f};
        if constexpr (CP) {
            y = __builtin_amdgcn_mfma_f32_16x16x4f32(la[0], x0.x, y, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_16x16x4f32(la[1], x1.x, y, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_16x16x4f32(la[2], x0.y, y, 0, 0, 0);
            y = __builtin_amdgcn_mfma_f32_16x16x4f32(la[3], x1.y, y, 0, 0, 0);
        }
        y = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], x0.x, y, 0, 0, 0);
        y = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], x1.x, y, 0, 0, 0);
        y = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], x0.y, y, 0, 0, 0);
        y = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], x1.y, y, 0, 0, 0);
        return y;
    };
    const int G = 1 << (m - __builtin_popcount(fixed));
    if (G >= 16 * kWaves) {
        // Every wave owns G / (16 kWaves) full column blocks: no lane
        // predicates and wave-uniform (scalar) block addresses from a lane
        // table.  Software-pipelined: block it's MFMA chain is issued, then
        // block it - 1's result (its chain finished meanwhile) is stored and
        // block it + 1's inputs are loaded, so a wave never waits on the
        // chain it has just issued.  A complex64 tile fits LDS up to 2^14
        // amplitudes, so a wave walks at most 32 column blocks and the table's
        // low half (`lo`) alone holds their slots.
        const int nit = G / (16 * kWaves);
        const BlockSlots bs = lw.on ? BlockSlots(BlockSlots::Given{}, blo, bhi) : BlockSlots(fixed, EB);
        int Uc = __builtin_amdgcn_readlane(bs.lo, 0), Up = Uc;
        float2 x0[NV], x1[NV];
        f32x4_t yp[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            x0[v] = lds_ld<float2>(lds + v * vstride, pb0 ^ Uc);
            x1[v] = lds_ld<float2>(lds + v * vstride, pb1 ^ Uc);
            yp[v] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        }
        for (int it = 0; it < nit; ++it) {
            f32x4_t y[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) y[v] = mv(x0[v], x1[v]);
            if (it > 0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    lds_st<float2>(lds + v * vstride, py0 ^ Up, make_float2(yp[v][0], yp[v][1]));
                    lds_st<float2>(lds + v * vstride, py1 ^ Up, make_float2(yp[v][2], yp[v][3]));
                }
            }
            const int Un = __builtin_amdgcn_readlane(bs.lo, (it + 1) & 63);
            if (it + 1 < nit) {
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    x0[v] = lds_ld<float2>(lds + v * vstride, pb0 ^ Un);
                    x1[v] = lds_ld<float2>(lds + v * vstride, pb1 ^ Un);
                }
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) yp[v] = y[v];
            Up = Uc;
            Uc = Un;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            lds_st<float2>(lds + v * vstride, py0 ^ Up, make_float2(yp[v][0], yp[v][1]));
            lds_st<float2>(lds + v * vstride, py1 ^ Up, make_float2(yp[v][2], yp[v][3]));
        }
        return;
    }
    const bool valid = i16 < G;  // tiles narrower than 16 columns
    const int step = (blockDim.x >> 6) * 16;
    constexpr int UN = 1;  // column blocks per iteration; 4 measured no faster than 1
    for (int c0 = (threadIdx.x >> 6) * 16; c0 < G; c0 += UN * step) {
        int U[UN];
        bool on[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            on[u] = c0 + u * step < G;  // uniform over the wave
            U[u] = on[u] ? slot(spread(c0 + u * step, fixed)) * EB : 0;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            lds_t *tb = lds + v * vstride;
            float2 x0[UN], x1[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                x0[u] = make_float2(0.f, 0.f);
                x1[u] = x0[u];
                if (valid && on[u]) {
                    x0[u] = lds_ld<float2>(tb, pb0 ^ U[u]);
                    x1[u] = lds_ld<float2>(tb, pb1 ^ U[u]);
                }
            }
            f32x4_t y[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) y[u] = mv(x0[u], x1[u]);
#pragma unroll
            for (int u = 0; u < UN; ++u)
                if (valid && on[u]) {
                    lds_st<float2>(tb, py0 ^ U[u], make_float2(y[u][0], y[u][1]));
                    lds_st<float2>(tb, py1 ^ U[u], make_float2(y[u][2], y[u][3]));
                }
        }
    }
}

// Diagonal op: matrix-index bit j is tile bit pos[j] (pos >= 0) or global bit
// -1-pos[j] outside the tile (constant over the tile, taken from `base`).
// Groups enumerate the tile bits that are neither local targets nor
// controls; each group holds 2^nl amplitudes with wave-uniform factors.
template <typename T, int NV, bool CP, typename MP>
__device__ inline void op_diag(lds_t *lds, int vstride, int m, const DevOp &op, MP Dg, MP DL, uint64_t base) {
    using C = typename VT<T>::c;
    constexpr int EB = sizeof(C);
    int gsub = 0, gmask = 0, lmask = 0;
    // fully unrolled: no run-time indexing of op.pos (keeps it out of scratch)
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        if (j >= op.k) break;
        const int pj = op.pos[j];
        if (pj < 0) {
            gmask |= 1 << j;
            if (base >> (-1 - pj) & 1) gsub |= 1 << j;
        } else {
            lmask |= 1 << pj;
        }
    }
    const int cl = (int)op.cloc;
    const Groups g(m, lmask | cl, EB);
    if (g.idle()) return;
    const int dim = 1 << op.k;
    for (int j = 0; j < g.J; ++j) {
        const int A = g.P ^ g.U(j, EB);
        // every matrix index d whose out-of-tile bits match this tile
        for (int d = 0; d < dim; ++d) {
            if ((d & gmask) != gsub) continue;
            int o = cl;
#pragma unroll
            for (int q = 0; q < 5; ++q)
                if (q < op.k && op.pos[q] >= 0 && (d >> q & 1)) o |= 1 << op.pos[q];
            o = slot(o) * EB;
            const C f = gload(Dg + d), fl = CP ? gload(DL + d) : f;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                lds_t *tb = lds + v * vstride;
                const C a = lds_ld<C>(tb, A ^ o);
                lds_st<C>(tb, A ^ o, CP ? cfma(f, a, cmul(fl, a)) : cmul(f, a));
            }
        }
    }
}

// Diagonal op with K = 1, 2 matrix bits: factors and slots of the entries
// present in this tile (those whose out-of-tile bits match `base`) are held in
// registers, and two groups are loaded before either is stored.
template <typename T, int NV, int K, bool CP, typename MP>
__device__ inline void op_diag_lo(lds_t *lds, int vstride, int m, const DevOp &op, MP Dg, MP DL, uint64_t base) {
    using C = typename VT<T>::c;
    constexpr int D = 1 << K;
    constexpr int EB = sizeof(C);
    const int cl = (int)op.cloc;
    int gsub = 0, gmask = 0, lmask = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int pj = op.pos[j];
        if (pj < 0) {
            gmask |= 1 << j;
            if (base >> (-1 - pj) & 1) gsub |= 1 << j;
        } else {
            lmask |= 1 << pj;
        }
    }
    const Groups g(m, lmask | cl, EB);
    if (g.idle()) return;
    bool on[D];
    int so[D];
    C f[D], fl[CP ? D : 1];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        on[s] = (s & gmask) == gsub;
        int o = cl;
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (op.pos[j] >= 0 && (s >> j & 1)) o |= 1 << op.pos[j];
        so[s] = slot(o) * EB;
        f[s] = gload(Dg + s);
        if constexpr (CP) fl[s] = gload(DL + s);
    }
    auto mul = [&](int s, C a) {
        if constexpr (CP) return cfma(f[s], a, cmul(fl[s], a));
        else return cmul(f[s], a);
    };
    for (int v = 0; v < NV; ++v) {  // not unrolled: one vector's registers at a time
        lds_t *tb = lds + v * vstride;
        int j = 0;
        for (; j + 2 <= g.J; j += 2) {
            const int A0 = g.P ^ g.U(j, EB), A1 = g.P ^ g.U(j + 1, EB);
            C a0[D], a1[D];
#pragma unroll
            for (int s = 0; s < D; ++s)
                if (on[s]) {
                    a0[s] = lds_ld<C>(tb, A0 ^ so[s]);
                    a1[s] = lds_ld<C>(tb, A1 ^ so[s]);
                }
#pragma unroll
            for (int s = 0; s < D; ++s)
                if (on[s]) {
                    lds_st<C>(tb, A0 ^ so[s], mul(s, a0[s]));
                    lds_st<C>(tb, A1 ^ so[s], mul(s, a1[s]));
                }
        }
        if (j < g.J) {
            const int A0 = g.P ^ g.U(j, EB);
#pragma unroll
            for (int s = 0; s < D; ++s)
                if (on[s]) lds_st<C>(tb, A0 ^ so[s], mul(s, lds_ld<C>(tb, A0 ^ so[s])));
        }
    }
}

// Global phase on the whole tile (order of elements irrelevant).
template <typename T, int NV, bool CP, typename MP>
__device__ inline void op_scalar(lds_t *lds, int vstride, int m, MP S, MP SL) {
    using C = typename VT<T>::c;
    const C f = gload(S), fl = CP ? gload(SL) : f;
    const int size = 1 << m;
    for (int i = threadIdx.x; i < size; i += blockDim.x)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            lds_t *tb = lds + v * vstride;
            const C x = lds_ld<C>(tb, i * (int)sizeof(C));
            lds_st<C>(tb, i * (int)sizeof(C), CP ? cfma(f, x, cmul(fl, x)) : cmul(f, x));
        }
}

// Re <lam| M |psi> over the tile (tile 0 = psi, tile 1 = lam); M is a dense
// 2x2 generator on tile bit pos[0] with local controls.  Complex64: each
// pair's product and inner product in FP32 (gate precision), the sum over the
// tile in FP64, two groups in flight.
template <typename MP>
__device__ inline double op_grad_dense_f32(lds_t *lds, int vstride, int m, const DevOp &op, MP M) {
    constexpr int EB = sizeof(float2);
    const float2 m00 = gload(M), m01 = gload(M + 1), m10 = gload(M + 2), m11 = gload(M + 3);
    const int cl = (int)op.cloc, tb = 1 << op.pos[0];
    const int o0 = slot(cl) * EB, o1 = slot(cl | tb) * EB;
    const Groups g(m, cl | tb, EB);
    if (g.idle()) return 0.0;
    const lds_t *ls = lds + vstride;
    auto pair = [&](int A) {
        const float2 s0 = lds_ld<float2>(lds, A ^ o0), s1 = lds_ld<float2>(lds, A ^ o1);
        const float2 l0 = lds_ld<float2>(ls, A ^ o0), l1 = lds_ld<float2>(ls, A ^ o1);
        const float2 g0 = cfma(m01, s1, cmul(m00, s0));
        const float2 g1 = cfma(m11, s1, cmul(m10, s0));
        return __builtin_fmaf(l1.y, g1.y, __builtin_fmaf(l1.x, g1.x, __builtin_fmaf(l0.y, g0.y, l0.x * g0.x)));
    };
    double acc = 0;
    int j = 0;
    for (; j + 2 <= g.J; j += 2) {
        const float a0 = pair(g.P ^ g.U(j, EB)), a1 = pair(g.P ^ g.U(j + 1, EB));
        acc += (double)a0 + (double)a1;
    }
    if (j < g.J) acc += (double)pair(g.P ^ g.U(j, EB));
    return acc;
}

template <typename MP>
__device__ inline double op_grad_diag_f32(lds_t *lds, int vstride, int m, const DevOp &op, MP Dg) {
    constexpr int EB = sizeof(float2);
    const float2 d0 = gload(Dg), d1 = gload(Dg + 1);
    const int cl = (int)op.cloc, tb = 1 << op.pos[0];
    const int o0 = slot(cl) * EB, o1 = slot(cl | tb) * EB;
    const Groups g(m, cl | tb, EB);
    if (g.idle()) return 0.0;
    const lds_t *ls = lds + vstride;
    auto pair = [&](int A) {
        const float2 g0 = cmul(d0, lds_ld<float2>(lds, A ^ o0)), g1 = cmul(d1, lds_ld<float2>(lds, A ^ o1));
        const float2 l0 = lds_ld<float2>(ls, A ^ o0), l1 = lds_ld<float2>(ls, A ^ o1);
        return __builtin_fmaf(l1.y, g1.y, __builtin_fmaf(l1.x, g1.x, __builtin_fmaf(l0.y, g0.y, l0.x * g0.x)));
    };
    double acc = 0;
    int j = 0;
    for (; j + 2 <= g.J; j += 2) {
        const float a0 = pair(g.P ^ g.U(j, EB)), a1 = pair(g.P ^ g.U(j + 1, EB));
        acc += (double)a0 + (double)a1;
    }
    if (j < g.J) acc += (double)pair(g.P ^ g.U(j, EB));
    return acc;
}

template <typename T, typename MP>
__device__ inline double op_grad_dense(lds_t *lds, int vstride, int m, const DevOp &op, MP M) {
    if constexpr (std::is_same<T, float>::value) return op_grad_dense_f32(lds, vstride, m, op, M);
    using C = typename VT<T>::c;
    constexpr int EB = sizeof(C);
    const double2 m00 = to_d(gload(M)), m01 = to_d(gload(M + 1)), m10 = to_d(gload(M + 2)),
                  m11 = to_d(gload(M + 3));
    const int cl = (int)op.cloc, tb = 1 << op.pos[0];
    const int o0 = slot(cl) * EB, o1 = slot(cl | tb) * EB;
    const Groups g(m, cl | tb, EB);
    if (g.idle()) return 0.0;
    const lds_t *ls = lds + vstride;
    double acc = 0;
    for (int j = 0; j < g.J; ++j) {
        const int A = g.P ^ g.U(j, EB);
        const double2 s0 = to_d(lds_ld<C>(lds, A ^ o0)), s1 = to_d(lds_ld<C>(lds, A ^ o1));
        const double2 l0 = to_d(lds_ld<C>(ls, A ^ o0)), l1 = to_d(lds_ld<C>(ls, A ^ o1));
        const double2 g0 = cfma(m01, s1, cmul(m00, s0));
        const double2 g1 = cfma(m11, s1, cmul(m10, s0));
        acc += l0.x * g0.x + l0.y * g0.y + l1.x * g1.x + l1.y * g1.y;
    }
    return acc;
}

// Same for a diagonal generator diag(d0, d1) on tile bit pos[0].
template <typename T, typename MP>
__device__ inline double op_grad_diag(lds_t *lds, int vstride, int m, const DevOp &op, MP Dg) {
    if constexpr (std::is_same<T, float>::value) return op_grad_diag_f32(lds, vstride, m, op, Dg);
    using C = typename VT<T>::c;
    constexpr int EB = sizeof(C);
    const double2 d0 = to_d(gload(Dg)), d1 = to_d(gload(Dg + 1));
    const int cl = (int)op.cloc, tb = 1 << op.pos[0];
    const int o0 = slot(cl) * EB, o1 = slot(cl | tb) * EB;
    const Groups g(m, cl | tb, EB);
    if (g.idle()) return 0.0;
    const lds_t *ls = lds + vstride;
    double acc = 0;
    for (int j = 0; j < g.J; ++j) {
        const int A = g.P ^ g.U(j, EB);
        const double2 g0 = cmul(d0, to_d(lds_ld<C>(lds, A ^ o0)));
        const double2 g1 = cmul(d1, to_d(lds_ld<C>(lds, A ^ o1)));
        const double2 l0 = to_d(lds_ld<C>(ls, A ^ o0)), l1 = to_d(lds_ld<C>(ls, A ^ o1));
        acc += l0.x * g0.x + l0.y * g0.y + l1.x * g1.x + l1.y * g1.y;
    }
    return acc;
}

// A gradient probe on tile bit p (generator G, dense 2x2 or diagonal)
// followed by a block of KB = 1, 2 bits (dense or diagonal, BD) that contains
// p, in one sweep: each group of the block is loaded once from both tiles,
// the probe's pairs along p are read off the loaded values, then the block is
// applied to both tiles.  Probe products in the state precision, the tile sum
// in FP64.
template <typename T, int KB, bool BD, bool CP, typename MP>
__device__ inline double op_grad_block_k(lds_t *lds, int vstride, int m, const DevOp &gop, MP G, const DevOp &bop,
                                         MP B, MP BL) {
    using C = typename VT<T>::c;
    using R = T;
    constexpr int D = 1 << KB;
    constexpr int EB = sizeof(C);
    const bool gdiag = gop.kind == OPK_GRADDIAG;
    const C g00 = gload(G), g11 = gdiag ? gload(G + 1) : gload(G + 3);
    const C g01 = gdiag ? make_zero<C>() : gload(G + 1), g10 = gdiag ? make_zero<C>() : gload(G + 2);
    const bool jp1 = KB == 2 && bop.pos[1] == gop.pos[0];  // probe on matrix bit 1
    int fixed = 0, so[D];
#pragma unroll
    for (int j = 0; j < KB; ++j) fixed |= 1 << bop.pos[j];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        int o = 0;
#pragma unroll
        for (int j = 0; j < KB; ++j)
            if (s >> j & 1) o |= 1 << bop.pos[j];
        so[s] = slot(o) * EB;
    }
    const Groups g(m, fixed, EB);
    if (g.idle()) return 0.0;
    C mm[BD ? D : D * D], ml[CP ? (BD ? D : D * D) : 1];
#pragma unroll
    for (int i = 0; i < (BD ? D : D * D); ++i) mm[i] = gload(B + i);
    if constexpr (CP) {
#pragma unroll
        for (int i = 0; i < (BD ? D : D * D); ++i) ml[i] = gload(BL + i);
    }
    // Re(conj(l0) (G s)_0 + conj(l1) (G s)_1) for the pair (0, 1) along p
    auto probe = [&](C s0, C s1, C l0, C l1) -> R {
        const C a0 = cfma(g01, s1, cmul(g00, s0)), a1 = cfma(g11, s1, cmul(g10, s0));
        return l0.x * a0.x + l0.y * a0.y + l1.x * a1.x + l1.y * a1.y;
    };
    lds_t *lt = lds + vstride;
    double acc = 0;
    for (int j = 0; j < g.J; ++j) {
        const int A = g.P ^ g.U(j, EB);
        C x[D], y[D];
#pragma unroll
        for (int s = 0; s < D; ++s) {
            x[s] = lds_ld<C>(lds, A ^ so[s]);
            y[s] = lds_ld<C>(lt, A ^ so[s]);
        }
        if constexpr (KB == 1) {
            acc += (double)probe(x[0], x[1], y[0], y[1]);
        } else if (jp1) {
            acc += (double)probe(x[0], x[2], y[0], y[2]) + (double)probe(x[1], x[3], y[1], y[3]);
        } else {
            acc += (double)probe(x[0], x[1], y[0], y[1]) + (double)probe(x[2], x[3], y[2], y[3]);
        }
#pragma unroll
        for (int r = 0; r < D; ++r) {
            C ox, oy;
            if constexpr (BD && CP) {
                ox = cfma(mm[r], x[r], cmul(ml[r], x[r]));
                oy = cfma(mm[r], y[r], cmul(ml[r], y[r]));
            } else if constexpr (BD) {
                ox = cmul(mm[r], x[r]);
                oy = cmul(mm[r], y[r]);
            } else if constexpr (CP) {
                ox = cmul(ml[r * D], x[0]);
                oy = cmul(ml[r * D], y[0]);
#pragma unroll
                for (int c = 1; c < D; ++c) {
                    ox = cfma(ml[r * D + c], x[c], ox);
                    oy = cfma(ml[r * D + c], y[c], oy);
                }
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    ox = cfma(mm[r * D + c], x[c], ox);
                    oy = cfma(mm[r * D + c], y[c], oy);
                }
            } else {
                ox = cmul(mm[r * D], x[0]);
                oy = cmul(mm[r * D], y[0]);
#pragma unroll
                for (int c = 1; c < D; ++c) {
                    ox = cfma(mm[r * D + c], x[c], ox);
                    oy = cfma(mm[r * D + c], y[c], oy);
                }
            }
            lds_st<C>(lds, A ^ so[r], ox);
            lds_st<C>(lt, A ^ so[r], oy);
        }
    }
    return acc;
}

template <typename T, bool CP, typename MP>
__device__ inline double op_grad_block(lds_t *lds, int vstride, int m, const DevOp &gop, MP G, const DevOp &bop,
                                       MP B, MP BL) {
    if (bop.kind == OPK_DIAG)
        return bop.k == 1 ? op_grad_block_k<T, 1, true, CP>(lds, vstride, m, gop, G, bop, B, BL)
                          : op_grad_block_k<T, 2, true, CP>(lds, vstride, m, gop, G, bop, B, BL);
    return bop.k == 1 ? op_grad_block_k<T, 1, false, CP>(lds, vstride, m, gop, G, bop, B, BL)
                      : op_grad_block_k<T, 2, false, CP>(lds, vstride, m, gop, G, bop, B, BL);
}

}  // namespace

// ---------------------------------------------------------------------------
// Register groups (complex64 forward stages of 2^12-amplitude tiles).
//
// Each wave holds 1024 amplitudes of the tile: lane l, register j (v[j]) holds
// tile index sum_i lane bit i << lpos[i] | sum_r (j >> r & 1) << rpos[r] |
// wave bits; a group reads them from LDS, exchanges lane and register bits in
// registers, applies its 3-bit blocks on MFMA and writes the tile back in the
// layout it ends in (the host tracks the layouts and emits the exchanges).
// ---------------------------------------------------------------------------

// In-place DPP exchange of one register pair (a: register bit clear, b: set;
// both floats) across lane bit B, as one asm block with two temporaries, so a
// large register-resident tile is updated in place.  Each DPP source written
// inside the block is written at least two instructions before it is read;
// the caller puts s_nop around the whole exchange for the sources written
// before it and for the MFMA reads after it.
template <int B> __device__ inline void rg_xchg(float2 &a, float2 &b) {
    float t0, t1;
    if constexpr (B == 2) {
        asm volatile("v_mov_b32 %4, %0\n\tv_mov_b32 %5, %1\n\t"
                     "v_mov_b32_dpp %0, %2 row_shr:4 row_mask:0xf bank_mask:0xa\n\t"
                     "v_mov_b32_dpp %1, %3 row_shr:4 row_mask:0xf bank_mask:0xa\n\t"
                     "v_mov_b32_dpp %2, %4 row_shl:4 row_mask:0xf bank_mask:0x5\n\t"
                     "v_mov_b32_dpp %3, %5 row_shl:4 row_mask:0xf bank_mask:0x5"
                     : "+v"(a.x), "+v"(a.y), "+v"(b.x), "+v"(b.y), "=&v"(t0), "=&v"(t1));
    } else if constexpr (B == 3) {
        asm volatile("v_mov_b32 %4, %0\n\tv_mov_b32 %5, %1\n\t"
                     "v_mov_b32_dpp %0, %2 row_shr:8 row_mask:0xf bank_mask:0xc\n\t"
                     "v_mov_b32_dpp %1, %3 row_shr:8 row_mask:0xf bank_mask:0xc\n\t"
                     "v_mov_b32_dpp %2, %4 row_shl:8 row_mask:0xf bank_mask:0x3\n\t"
                     "v_mov_b32_dpp %3, %5 row_shl:8 row_mask:0xf bank_mask:0x3"
                     : "+v"(a.x), "+v"(a.y), "+v"(b.x), "+v"(b.y), "=&v"(t0), "=&v"(t1));
    } else if constexpr (B == 1) {
        // vcc = lanes with lane bit 1 set: b' = vcc ? b : a[l ^ 2]; then
        // vcc = clear: a' = vcc ? a : b[l ^ 2]
        asm volatile("v_mov_b32 %4, %2\n\tv_mov_b32 %5, %3\n\t"
                     "s_mov_b32 vcc_lo, 0xcccccccc\n\ts_mov_b32 vcc_hi, 0xcccccccc\n\t"
                     "v_cndmask_b32_dpp %2, %0, %2, vcc quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                     "v_cndmask_b32_dpp %3, %1, %3, vcc quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                     "s_not_b64 vcc, vcc\n\t"
                     "v_cndmask_b32_dpp %0, %4, %0, vcc quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
                     "v_cndmask_b32_dpp %1, %5, %1, vcc quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf"
                     : "+v"(a.x), "+v"(a.y), "+v"(b.x), "+v"(b.y), "=&v"(t0), "=&v"(t1)
                     :
                     : "vcc");
    } else {
        asm volatile("v_mov_b32 %4, %2\n\tv_mov_b32 %5, %3\n\t"
                     "s_mov_b32 vcc_lo, 0xaaaaaaaa\n\ts_mov_b32 vcc_hi, 0xaaaaaaaa\n\t"
                     "v_cndmask_b32_dpp %2, %0, %2, vcc quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                     "v_cndmask_b32_dpp %3, %1, %3, vcc quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                     "s_not_b64 vcc, vcc\n\t"
                     "v_cndmask_b32_dpp %0, %4, %0, vcc quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
                     "v_cndmask_b32_dpp %1, %5, %1, vcc quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf"
                     : "+v"(a.x), "+v"(a.y), "+v"(b.x), "+v"(b.y), "=&v"(t0), "=&v"(t1)
                     :
                     : "vcc");
    }
}

// Exchange register bit R with lane bit B: the amplitude at (lane bit B = x,
// register bit R = y) moves to (lane bit B = y, register bit R = x).  For
// each register pair (a = v[j], b = v[j | 2^R]) lanes with bit B set take
// b's value from lane l ^ 2^B into a, lanes with it clear take a's value from
// lane l ^ 2^B into b.  Lane bits 4, 5: one v_permlane16/32_swap per float;
// lane bits 2, 3: DPP row shifts with bank masks; lane bits 0, 1: quad
// permutations in a masked select.
template <int R, int B, int NJ> __device__ inline void rg_xlane(float2 (&v)[NJ]) {
    if constexpr (B < 4) asm volatile("s_nop 1");
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        if (j >> R & 1) continue;
        float2 &a = v[j];
        float2 &b = v[j | (1 << R)];
        if constexpr (B == 5 || B == 4) {
            uint32_t ax = __float_as_uint(a.x), ay = __float_as_uint(a.y);
            uint32_t bx = __float_as_uint(b.x), by = __float_as_uint(b.y);
            if constexpr (B == 5) {
                const auto x = __builtin_amdgcn_permlane32_swap(ax, bx, false, false);
                const auto y = __builtin_amdgcn_permlane32_swap(ay, by, false, false);
                ax = x[0], bx = x[1], ay = y[0], by = y[1];
            } else {
                const auto x = __builtin_amdgcn_permlane16_swap(ax, bx, false, false);
                const auto y = __builtin_amdgcn_permlane16_swap(ay, by, false, false);
                ax = x[0], bx = x[1], ay = y[0], by = y[1];
            }
            a = make_float2(__uint_as_float(ax), __uint_as_float(ay));
            b = make_float2(__uint_as_float(bx), __uint_as_float(by));
        } else {
            rg_xchg<B>(a, b);
        }
    }
    if constexpr (B < 4) asm volatile("s_nop 4");
}

// One 3-bit block: targets at lane bits 4, 5 (the MFMA row groups) and
// register bit S; columns are lane bits 0..3 and the other register bits.
// Register pair (v[j], v[j | 2^S]) is the B operand [Re x0, Re x1, Im x0,
// Im x1] of one stacked real 16x16 product; the host orders the matrix rows
// so the accumulator comes back as (Re y0, Im y0, Re y1, Im y1) in the same
// layout.  Compensated: lo fragments accumulated first.
template <int S, bool CP, int NJ>
__device__ inline void rg_mfma3(float2 (&v)[NJ], const float4 &fa, const float4 &fl) {
    // the pairs in batches of IL independent accumulator chains, issued
    // interleaved so a wave has IL MFMAs in flight instead of waiting on each
    // chain's previous product
    constexpr int NP = NJ / 2, IL = NP >= 32 ? 2 : NP >= 4 ? 4 : NP;
    auto pj = [](int p) { return ((p >> S) << (S + 1)) | (p & ((1 << S) - 1)); };  // p-th pair's low register
#pragma unroll
    for (int p0 = 0; p0 < NP; p0 += IL) {
        f32x4_t y[IL];
#pragma unroll
        for (int i = 0; i < IL; ++i) y[i] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        if constexpr (CP) {
#pragma unroll
            for (int i = 0; i < IL; ++i) y[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(fl.x, v[pj(p0 + i)].x, y[i], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < IL; ++i)
                y[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(fl.y, v[pj(p0 + i) | (1 << S)].x, y[i], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < IL; ++i) y[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(fl.z, v[pj(p0 + i)].y, y[i], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < IL; ++i)
                y[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(fl.w, v[pj(p0 + i) | (1 << S)].y, y[i], 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < IL; ++i) y[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa.x, v[pj(p0 + i)].x, y[i], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < IL; ++i)
            y[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa.y, v[pj(p0 + i) | (1 << S)].x, y[i], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < IL; ++i) y[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa.z, v[pj(p0 + i)].y, y[i], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < IL; ++i)
            y[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa.w, v[pj(p0 + i) | (1 << S)].y, y[i], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < IL; ++i) {
            v[pj(p0 + i)] = make_float2(y[i][0], y[i][1]);
            v[pj(p0 + i) | (1 << S)] = make_float2(y[i][2], y[i][3]);
        }
    }
}

// Runs micro-ops mu[0 .. nmu) on the lane's registers v[0 .. NJ) (NJ = 2^NR
// amplitudes, NR register bits).  `q`: index of the next MFMA op of the
// stage, whose fragments `na` (hi), `nl` (lo) are already requested; the
// following op's fragments are requested while an op runs.
template <bool CP, int NJ>
__device__ inline void rg_run_micro(float2 (&v)[NJ], const DESV_CONST int32_t *mu, int nmu,
                                    const DESV_GLOBAL float4 *fr, int &q, int nq, float4 &na, float4 &nl) {
    constexpr int NR = NJ == 16 ? 4 : NJ == 32 ? 5 : 6;
    static_assert((1 << NR) == NJ, "register bits");
    constexpr int FB = CP ? 2 : 1;  // fragment blocks per MFMA op
    const int lane = threadIdx.x & 63;
    for (int u = 0; u < nmu; ++u) {
        const int code = mu[u];
        const int a1 = (code >> 4) & 15;
        if ((code & 15) == RG_XLANE) {
            const int b = (code >> 8) & 15;
            switch (a1 * 8 + b) {
#define DESV_RX(R)                                             \
    case R * 8 + 0: rg_xlane<R, 0>(v); break;                  \
    case R * 8 + 1: rg_xlane<R, 1>(v); break;                  \
    case R * 8 + 2: rg_xlane<R, 2>(v); break;                  \
    case R * 8 + 3: rg_xlane<R, 3>(v); break;                  \
    case R * 8 + 4: rg_xlane<R, 4>(v); break;                  \
    case R * 8 + 5: rg_xlane<R, 5>(v); break;
                DESV_RX(0) DESV_RX(1) DESV_RX(2) DESV_RX(3)
#undef DESV_RX
                default:
                    if constexpr (NR >= 5) {
                        switch (a1 * 8 + b) {
#define DESV_RX(R)                                             \
    case R * 8 + 0: rg_xlane<R, 0>(v); break;                  \
    case R * 8 + 1: rg_xlane<R, 1>(v); break;                  \
    case R * 8 + 2: rg_xlane<R, 2>(v); break;                  \
    case R * 8 + 3: rg_xlane<R, 3>(v); break;                  \
    case R * 8 + 4: rg_xlane<R, 4>(v); break;                  \
    case R * 8 + 5: rg_xlane<R, 5>(v); break;
                            DESV_RX(4)
                            default:
                                if constexpr (NR >= 6) {
                                    switch (a1 * 8 + b) {
                                        DESV_RX(5)
                                        default: break;
                                    }
                                }
                                break;
#undef DESV_RX
                        }
                    }
                    break;
            }
        } else {
            const float4 fa = na, fl = nl;
            if (q + 1 < nq) {
                // wave-uniform base + 32-bit lane offset: no 64-bit per-lane
                // address held across the micro-op loop
                const uint32_t i0 = (uint32_t)((q + 1) * FB * 64) + (uint32_t)lane;
                na = gload(fr + i0);
                if constexpr (CP) nl = gload(fr + (i0 + 64u));
            }
            ++q;
            switch (a1) {
                case 0: rg_mfma3<0, CP>(v, fa, fl); break;
                case 1: rg_mfma3<1, CP>(v, fa, fl); break;
                case 2: rg_mfma3<2, CP>(v, fa, fl); break;
                case 3: rg_mfma3<3, CP>(v, fa, fl); break;
                default:
                    if constexpr (NR >= 5) {
                        if (a1 == 4) rg_mfma3<4, CP>(v, fa, fl);
                        if constexpr (NR >= 6)
                            if (a1 == 5) rg_mfma3<5, CP>(v, fa, fl);
                    }
                    break;
            }
        }
    }
}

// byte slot of this lane's part of a layout (lane and wave bits)
__device__ inline int rg_lane_slot(const DESV_CONST int32_t *lpos, const DESV_CONST int32_t *wpos) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int idx = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) idx |= ((lane >> i) & 1) << lpos[i];
#pragma unroll
    for (int i = 0; i < 2; ++i) idx |= ((wave >> i) & 1) << wpos[i];
    return slot(idx) * (int)sizeof(float2);
}

// `na`, `nl`: the first MFMA op's fragments (hi, lo), already requested.
template <bool CP>
__device__ inline void run_reg_groups(lds_t *lds, const DESV_CONST int32_t *rp, const DESV_GLOBAL float4 *fr,
                                      float4 na, float4 nl) {
    const int ng = rp[0], nq = rp[1];
    int q = 0, w = 2;
    float2 v[16];
    for (int g = 0; g < ng; ++g) {
        const DESV_CONST int32_t *h = rp + w;
        const int nmu = h[0];
        {
            const int lb = rg_lane_slot(h + 1, h + 7);
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = lds_ld<float2>(lds, lb ^ h[9 + j]);
        }
        rg_run_micro<CP>(v, h + kRegGroupHeader, nmu, fr, q, nq, na, nl);
        {
            const int lb = rg_lane_slot(h + 25, h + 7);
#pragma unroll
            for (int j = 0; j < 16; ++j) lds_st<float2>(lds, lb ^ h[31 + j], v[j]);
        }
        __syncthreads();
        w += kRegGroupHeader + nmu;
    }
}

// One 3-qubit block of a wave tile in the 4-block MFMA form: targets on
// register bits X < Y < Z; for each group of 8 amplitudes (the other
// register bits), 16 k-steps of v_mfma_f32_16x16x1_4b_f32, k-step k reading
// float (k & 1) of the register with bits X, Y, Z = k >> 1 as the lane's B
// value and the fragment a[k] (row lane % 16 of the stacked real matrix) as
// its A value; accumulator register r goes back to the register k = r came
// from.  The instruction returns input lane group b's outputs in
// accumulator registers 4b .. 4b + 3 and output rows 4g' .. 4g' + 3 in lane
// group g', so the op also exchanges lane bits 4, 5 with register bits Y, Z
// (the host plans for it).  IL groups are in flight at a time.
typedef float f32x16_t __attribute__((ext_vector_type(16)));
// Build knobs of the wave-tile kernel: MFMA chains in flight per wave (1: 166
// VGPRs at 64 amplitudes per lane; 2 needs 256 and spills) and waves per SIMD
// at 32 / 64 amplitudes per lane.
#ifndef DESV_WT_IL
#define DESV_WT_IL 1
#endif
#ifndef DESV_WT_W5
#define DESV_WT_W5 3
#endif
#ifndef DESV_WT_W6
#define DESV_WT_W6 3
#endif
// The accumulator's 8 register pairs into the tile registers o0..o7, one
// 64-bit v_pk_mov_b32 each, in one asm block: the compiler neither merges the
// tile registers into the accumulators' 16-register tuples (whose grouping
// differs from block to block; it spilled when it did) nor tracks MFMA
// hazards into inline asm, so the block waits out the MFMA result latency
// itself (16 wait states before the first read) and leaves 4 after its last
// write for the MFMA, permlane or DPP reads that follow.
__device__ inline void rg_copy8(float2 &o0, float2 &o1, float2 &o2, float2 &o3, float2 &o4, float2 &o5,
                                float2 &o6, float2 &o7, const f32x16_t &d) {
    const float2 p0 = make_float2(d[0], d[1]), p1 = make_float2(d[2], d[3]), p2 = make_float2(d[4], d[5]),
                 p3 = make_float2(d[6], d[7]), p4 = make_float2(d[8], d[9]), p5 = make_float2(d[10], d[11]),
                 p6 = make_float2(d[12], d[13]), p7 = make_float2(d[14], d[15]);
    asm volatile("s_nop 7\n\ts_nop 7\n\t"
                 "v_pk_mov_b32 %0, %8, %8 op_sel:[0,1]\n\t"
                 "v_pk_mov_b32 %1, %9, %9 op_sel:[0,1]\n\t"
                 "v_pk_mov_b32 %2, %10, %10 op_sel:[0,1]\n\t"
                 "v_pk_mov_b32 %3, %11, %11 op_sel:[0,1]\n\t"
                 "v_pk_mov_b32 %4, %12, %12 op_sel:[0,1]\n\t"
                 "v_pk_mov_b32 %5, %13, %13 op_sel:[0,1]\n\t"
                 "v_pk_mov_b32 %6, %14, %14 op_sel:[0,1]\n\t"
                 "v_pk_mov_b32 %7, %15, %15 op_sel:[0,1]\n\t"
                 "s_nop 3"
                 : "=&v"(o0), "=&v"(o1), "=&v"(o2), "=&v"(o3), "=&v"(o4), "=&v"(o5), "=&v"(o6), "=&v"(o7)
                 : "v"(p0), "v"(p1), "v"(p2), "v"(p3), "v"(p4), "v"(p5), "v"(p6), "v"(p7));
}
template <int X, int Y, int Z, int NR>
__device__ constexpr int rg_treg(int g, int t) {
    int j = 0, gi = 0;
    for (int b = 0; b < NR; ++b) {
        if (b == X) j |= (t & 1) << b;
        else if (b == Y) j |= ((t >> 1) & 1) << b;
        else if (b == Z) j |= ((t >> 2) & 1) << b;
        else j |= ((g >> gi++) & 1) << b;
    }
    return j;
}
__device__ inline float &rg_f(float2 &x, int c) { return c ? x.y : x.x; }
// 16 A values (one fragment row) from LDS
__device__ inline void rg_frag_row(const lds_t *f, float (&a)[16]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float4 x = lds_ld<float4>(f, 16 * c);
        a[4 * c] = x.x, a[4 * c + 1] = x.y, a[4 * c + 2] = x.z, a[4 * c + 3] = x.w;
    }
}
// Compensated blocks (CP) read their hi and lo rows from LDS (`f`, the lo
// block 1 KiB after the hi one) for every group instead of holding both, so
// the 64-amplitude tile, one accumulator and one row fit 3 waves per SIMD.
template <int X, int Y, int Z, bool CP, int NJ>
__device__ inline void rg_mfma3t(float2 (&v)[NJ], const float (&a0)[16], const lds_t *f) {
    constexpr int NR = NJ == 32 ? 5 : 6, NG = NJ / 8, IL = DESV_WT_IL;
#pragma unroll
    for (int g0 = 0; g0 < NG; g0 += IL) {
        f32x16_t d[IL];
#pragma unroll
        for (int i = 0; i < IL; ++i) d[i] = f32x16_t{};
        float a[16];
        if constexpr (CP) {
            float al[16];
            rg_frag_row(f + 1024, al);
#pragma unroll
            for (int k = 0; k < 16; ++k)
#pragma unroll
                for (int i = 0; i < IL; ++i)
                    d[i] = __builtin_amdgcn_mfma_f32_16x16x1f32(al[k], rg_f(v[rg_treg<X, Y, Z, NR>(g0 + i, k >> 1)], k & 1),
                                                                d[i], 0, 0, 0);
            rg_frag_row(f, a);
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = a0[k];
        }
#pragma unroll
        for (int k = 0; k < 16; ++k)
#pragma unroll
            for (int i = 0; i < IL; ++i)
                d[i] = __builtin_amdgcn_mfma_f32_16x16x1f32(a[k], rg_f(v[rg_treg<X, Y, Z, NR>(g0 + i, k >> 1)], k & 1),
                                                            d[i], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < IL; ++i) {
            rg_copy8(v[rg_treg<X, Y, Z, NR>(g0 + i, 0)], v[rg_treg<X, Y, Z, NR>(g0 + i, 1)],
                     v[rg_treg<X, Y, Z, NR>(g0 + i, 2)], v[rg_treg<X, Y, Z, NR>(g0 + i, 3)],
                     v[rg_treg<X, Y, Z, NR>(g0 + i, 4)], v[rg_treg<X, Y, Z, NR>(g0 + i, 5)],
                     v[rg_treg<X, Y, Z, NR>(g0 + i, 6)], v[rg_treg<X, Y, Z, NR>(g0 + i, 7)], d[i]);
        }
    }
}

// Micro-ops of a wave tile: exchanges and 4-block MFMA blocks.  `fl` = the
// stage's fragment blocks in LDS (FB = 1 or 2 blocks of 16 x 16 floats per
// op); a lane reads row lane % 16 of each.
template <bool CP, int NJ>
__device__ inline void rg_run_wave(float2 (&v)[NJ], const DESV_CONST int32_t *mu, int nmu, const lds_t *fl) {
    constexpr int NR = NJ == 32 ? 5 : 6, FB = CP ? 2 : 1;
    const int row = (threadIdx.x & 15) * 16 * (int)sizeof(float);
    int q = 0;
    for (int u = 0; u < nmu; ++u) {
        const int code = mu[u];
        if ((code & 15) == RG_XLANE) {
            const int a1 = (code >> 4) & 15, b = (code >> 8) & 15;
            switch (a1 * 8 + b) {
#define DESV_RX(R)                                             \
    case R * 8 + 0: rg_xlane<R, 0>(v); break;                  \
    case R * 8 + 1: rg_xlane<R, 1>(v); break;                  \
    case R * 8 + 2: rg_xlane<R, 2>(v); break;                  \
    case R * 8 + 3: rg_xlane<R, 3>(v); break;                  \
    case R * 8 + 4: rg_xlane<R, 4>(v); break;                  \
    case R * 8 + 5: rg_xlane<R, 5>(v); break;
                DESV_RX(0) DESV_RX(1) DESV_RX(2) DESV_RX(3) DESV_RX(4)
                default:
                    if constexpr (NR >= 6) {
                        switch (a1 * 8 + b) {
                            DESV_RX(5)
                            default: break;
                        }
                    }
                    break;
#undef DESV_RX
            }
        } else {
            float a[16];
            const lds_t *f = fl + (size_t)q * FB * 1024 + row;
            if constexpr (!CP) rg_frag_row(f, a);
            ++q;
            switch (code >> 4) {
#define DESV_RT(X, Y, Z) \
    case X | Y << 4 | Z << 8: rg_mfma3t<X, Y, Z, CP>(v, a, f); break;
                DESV_RT(0, 1, 2) DESV_RT(0, 1, 3) DESV_RT(0, 1, 4) DESV_RT(0, 2, 3) DESV_RT(0, 2, 4)
                DESV_RT(0, 3, 4) DESV_RT(1, 2, 3) DESV_RT(1, 2, 4) DESV_RT(1, 3, 4) DESV_RT(2, 3, 4)
                default:
                    if constexpr (NR >= 6) {
                        switch (code >> 4) {
                            DESV_RT(0, 1, 5) DESV_RT(0, 2, 5) DESV_RT(0, 3, 5) DESV_RT(0, 4, 5) DESV_RT(1, 2, 5)
                            DESV_RT(1, 3, 5) DESV_RT(1, 4, 5) DESV_RT(2, 3, 5) DESV_RT(2, 4, 5) DESV_RT(3, 4, 5)
                            default: break;
                        }
                    }
                    break;
#undef DESV_RT
            }
        }
    }
}

// Wave-tile stage: each of the 4 waves of a workgroup holds one tile of 2^m
// amplitudes (m = 6 + NR: 64 lanes x 2^NR registers) from its HBM load to its
// HBM store, with no barrier after the workgroup has copied the stage's
// fragment blocks into LDS, so a CU's waves drift apart and one wave's HBM
// traffic overlaps other waves' MFMA work.  The load layout puts tile bit 0
// at register bit 0 and tile bits 1..4 at lane bits 0..3 (each 16-byte load
// is an amplitude pair, 16 lanes read 256 contiguous bytes), the store layout
// does the same for the store order; the host's exchanges and the 4-block
// MFMA form move bits between.  A relabelling stage (flag word after the
// micro-ops) stores into tk.sv1 with its outer bits relabelled.
template <int NR, bool CP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NR <= 5 ? DESV_WT_W5 : DESV_WT_W6)))
k_wstage(StageTask tk) {
    constexpr int NJ = 1 << NR, M = 6 + NR;
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t zero = 0;
    lds_t *lds = reinterpret_cast<lds_t *>((uintptr_t)zero);
    const int lane = threadIdx.x & 63;
    // the wave's tile number, wave-uniform (kept in scalar registers)
    const int64_t t = (int64_t)uni64((uint64_t)((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)));
    const bool live = t < uni64((uint64_t)tk.st.n_tiles);
    const DESV_CONST int32_t *wp = cptr(tk.rprog);
    const int nmu = wp[0], nq = wp[1];
    // tile base: the tile number's bits deposited into the non-tile bits
    const uint64_t tile_mask = uni64(tk.st.tile_mask);
    const int n = uni(tk.n);
    const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    uint64_t base = 0;
    {
        uint64_t tt = (uint64_t)t, fm = ~tile_mask & all;
        while (fm) {
            const uint64_t lb = fm & (~fm + 1);
            if (tt & 1) base |= lb;
            tt >>= 1;
            fm ^= lb;
        }
    }
    DESV_GLOBAL float4 *sv = gptr_rw(reinterpret_cast<float4 *>(tk.sv0));
    // amplitude offset of this lane (lane bits) and of register pair p
    // (register bits 1.., wave-uniform) in a layout given as global bits
    auto lane_off = [&](const DESV_CONST int32_t *gb) {
        uint64_t o = base;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            if ((lane >> i) & 1) o |= 1ull << gb[i];
        return o;
    };
    auto reg_off = [&](const DESV_CONST int32_t *gb, int p) {
        uint64_t o = 0;
#pragma unroll
        for (int r = 1; r < NR; ++r)
            if ((p >> (r - 1)) & 1) o |= 1ull << gb[6 + r];
        return o;
    };
    float2 v[NJ];
    if (live) {
        const DESV_CONST int32_t *gb = wp + 2;
        const uint64_t lo = lane_off(gb);
#pragma unroll
        for (int p = 0; p < NJ / 2; ++p) {
            const float4 x = gload(sv + ((lo | reg_off(gb, p)) >> 1));
            v[2 * p] = make_float2(x.x, x.y);
            v[2 * p + 1] = make_float2(x.z, x.w);
        }
    }
    {
        // the stage's fragment blocks into LDS, 4 float4s a thread at a time;
        // the loads are unconditional (indices clamped into the table), so
        // the wait before each store counts only the loads in flight instead
        // of draining the tile loads issued above at every step
        const DESV_GLOBAL float4 *fr = gptr(reinterpret_cast<const float4 *>(tk.rfrag));
        const int cnt = nq * (CP ? 2 : 1) * 64;
        for (int i0 = 0; i0 < cnt; i0 += 4 * 256) {
            float4 x[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int i = i0 + c * 256 + (int)threadIdx.x;
                x[c] = gload(fr + (i < cnt ? i : cnt - 1));
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int i = i0 + c * 256 + (int)threadIdx.x;
                if (i < cnt) lds_st<float4>(lds, i * 16, x[c]);
            }
        }
    }
    __syncthreads();
    if (!live) return;
    rg_run_wave<CP>(v, wp + 2 + 2 * M, nmu, lds);
    {
        // a relabelling stage writes out of place, its outer bits moved to
        // the listed destination bits
        const DESV_CONST int32_t *rl = wp + 2 + 2 * M + nmu;
        DESV_GLOBAL float4 *dst = sv;
        if (rl[0]) {
            dst = gptr_rw(reinterpret_cast<float4 *>(tk.sv1));
            uint64_t tt = (uint64_t)t;
            base = 0;
            for (int i = 0; i < n - M; ++i, tt >>= 1)
                if (tt & 1) base |= 1ull << rl[1 + i];
        }
        const DESV_CONST int32_t *gb = wp + 2 + M;
        const uint64_t lo = lane_off(gb);
#pragma unroll
        for (int p = 0; p < NJ / 2; ++p)
            gstore(dst + ((lo | reg_off(gb, p)) >> 1), make_float4(v[2 * p].x, v[2 * p].y, v[2 * p + 1].x, v[2 * p + 1].y));
    }
}

hipError_t launch_wave_stage(int m, bool comp, int nq, const StageTask &t, hipStream_t s) {
    if (t.st.n_tiles <= 0) return hipSuccess;
    if (t.st.n_tiles > 0x1ffffffffll) return hipErrorInvalidValue;
    if (nq < 0 || nq * (comp ? 2 : 1) > kWaveMaxFragBlocks) return hipErrorInvalidValue;
    const dim3 g((unsigned)((t.st.n_tiles + 3) / 4));
    const size_t lds = (size_t)nq * (comp ? 2 : 1) * 1024;
    if (m == 11) {
        if (comp) hipLaunchKernelGGL((k_wstage<5, true>), g, dim3(256), lds, s, t);
        else hipLaunchKernelGGL((k_wstage<5, false>), g, dim3(256), lds, s, t);
    } else if (m == 12) {
        if (comp) hipLaunchKernelGGL((k_wstage<6, true>), g, dim3(256), lds, s, t);
        else hipLaunchKernelGGL((k_wstage<6, false>), g, dim3(256), lds, s, t);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stage kernel
// ---------------------------------------------------------------------------
// VAR selects the compiled micro-ops (register budget and occupancy):
//   0  1-2 qubit dense, diagonal, scalar and gradient ops;
//   1  + k = 3..5 VALU blocks;
//   2  + k = 3, 4 MFMA blocks (complex64; k = 3 VALU blocks kept as well);
//   3  + k = 5 MFMA blocks (their fragments double the register budget).
//   CP compensated matrices (lo parts in registers beside the hi parts; with
//   k = 5 MFMA blocks they need more than the 128 VGPRs of 4 waves per SIMD).
template <typename T, int NV, int VAR, bool CP>
__global__ void __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(std::is_same<T, float>::value ? (CP && VAR == 3 ? 2 : 4) : 1)))
k_stage(const StageTask *__restrict__ tasks, const int32_t *__restrict__ wg_task, StageTask single, int xoff) {
    using C = typename VT<T>::c;
    using V = typename VT<T>::v;
    constexpr int PV = VT<T>::per_vec;
    constexpr int EB = sizeof(C);
    constexpr int CH = kThreads * PV;  // amplitudes per load/store chunk
    // All of the kernel's LDS is dynamic, so the tile(s) start at address 0
    // (a slot is the address: no base add per access) and the bookkeeping
    // sits behind them, at byte xoff (kStageLdsExtra bytes).
    extern __shared__ __align__(16) unsigned char smem[];
    StageTask &tk = *reinterpret_cast<StageTask *>(smem + xoff);
    uint64_t *jtab = reinterpret_cast<uint64_t *>(smem + xoff + kStageLdsTask);       // global offset of chunk j
    int *qtab = reinterpret_cast<int *>(smem + xoff + kStageLdsTask + 256 * 8);       // byte slot of the tile index stored at chunk j
    double *gacc = reinterpret_cast<double *>(smem + xoff + kStageLdsTask + 256 * 12);

    {
        // copy the task into LDS in 4-byte words, one word per thread
        constexpr int W = (int)(sizeof(StageTask) / 4);
        const uint32_t *src = tasks ? reinterpret_cast<const uint32_t *>(tasks + wg_task[blockIdx.x])
                                    : reinterpret_cast<const uint32_t *>(&single);
        for (int i = threadIdx.x; i < W; i += blockDim.x) reinterpret_cast<uint32_t *>(&tk)[i] = src[i];
    }
    __syncthreads();
    const int64_t t = (int64_t)blockIdx.x - (int64_t)uni64((uint64_t)tk.wg_begin);
    if (t < 0 || t >= (int64_t)uni64((uint64_t)tk.st.n_tiles)) return;  // uniform over the workgroup

    const int m = uni(tk.st.m), L = uni(tk.st.L);
    const int size = 1 << m;
    // the tile at address 0 (an integer, not a null-pointer constant)
    const uint32_t zero = 0;
    lds_t *lds = reinterpret_cast<lds_t *>((uintptr_t)zero);
    const int vstride = size * EB;
    DESV_GLOBAL C *svp[2] = {gptr_rw(reinterpret_cast<C *>(tk.sv0)), gptr_rw(reinterpret_cast<C *>(tk.sv1))};
    const uint64_t tile_mask = uni64(tk.st.tile_mask);
    const int n = uni(tk.n);
    const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1);
    const DESV_GLOBAL C *mats = gptr(reinterpret_cast<const C *>(tk.mats));
    // lo halves of a compensated program's matrices: the pool's second half
    const DESV_GLOBAL C *matl = CP ? mats + uni(tk.lo_off) : mats;
    const DESV_CONST DevOp *ops = cptr(tk.ops);
    const DESV_GLOBAL int4 *ltab = tk.ltab ? gptr(reinterpret_cast<const int4 *>(tk.ltab)) : nullptr;
    const bool permuted = uni(tk.st.permuted) != 0;
    const int n_grad = uni(tk.st.n_grad);

    // global offset of tile-local index e (bits < L stay, bit L+j -> hbits[j])
    auto goff = [&](int e) -> uint64_t {
        uint64_t o = (uint64_t)(e & ((1 << L) - 1));
        for (int r = e >> L, j = 0; r; r >>= 1, ++j)
            if (r & 1) o |= 1ull << tk.st.hbits[j];
        return o;
    };
    // tile-local index stored at store-order index e (write-back permutation)
    auto perm = [&](int e) -> int {
        int o = 0;
        for (int j = 0; e; e >>= 1, ++j)
            if (e & 1) o |= 1 << tk.st.sperm[j];
        return o;
    };
    const int nchunks = size >= CH ? size / CH : 1;
    for (int i = threadIdx.x; i < n_grad * (kThreads / 64); i += blockDim.x) gacc[i] = 0.0;
    for (int j = threadIdx.x; j < nchunks; j += blockDim.x) {
        jtab[j] = goff(j * CH);
        qtab[j] = slot(permuted ? perm(j * CH) : j * CH) * EB;
    }
    // tile base: the tile number's bits deposited into the non-tile bits
    uint64_t base = 0;
    {
        uint64_t tt = (uint64_t)t, fm = ~tile_mask & all;
        while (fm) {
            const uint64_t lb = fm & (~fm + 1);
            if (tt & 1) base |= lb;
            tt >>= 1;
            fm ^= lb;
        }
    }
    const int le = (int)threadIdx.x * PV;  // this lane's first amplitude of a chunk
    const bool lane_on = le < size;
    const uint64_t gl = lane_on ? base | goff(le) : 0;
    const int sl = slot(le) * EB;
    // register groups: the first MFMA op's fragments, requested before the tile loads
    float4 rfa = make_float4(0.f, 0.f, 0.f, 0.f), rfl = rfa;
    const DESV_CONST int32_t *rprog = nullptr;
    const DESV_GLOBAL float4 *rfrag = nullptr;
    if constexpr (VAR == 4) {
        rprog = cptr(tk.rprog);
        rfrag = gptr(reinterpret_cast<const float4 *>(tk.rfrag));
        if (rprog[1] > 0) {
            rfa = gload(rfrag + (threadIdx.x & 63));
            if constexpr (CP) rfl = gload(rfrag + 64 + (threadIdx.x & 63));
        }
    }
    int ql[PV];
#pragma unroll
    for (int i = 0; i < PV; ++i) ql[i] = slot(permuted ? perm(le + i) : le + i) * EB;
    __syncthreads();  // jtab / qtab / gacc ready

    // load: chunk j covers amplitudes j*CH .. j*CH + CH - 1 of the tile; LB
    // chunks of loads in flight before their LDS writes
    constexpr int LB = 8 / NV;
    if (lane_on)
        for (int jb = 0; jb < nchunks; jb += LB) {
            V x[LB][NV];
#pragma unroll
            for (int u = 0; u < LB; ++u)
                if (jb + u < nchunks) {
                    const uint64_t off = gl | jtab[jb + u];
#pragma unroll
                    for (int v = 0; v < NV; ++v) x[u][v] = gload((const DESV_GLOBAL V *)(svp[v] + off));
                }
#pragma unroll
            for (int u = 0; u < LB; ++u)
                if (jb + u < nchunks) {
                    const int S = slot((jb + u) * CH) * EB;
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const C *xc = reinterpret_cast<const C *>(&x[u][v]);
#pragma unroll
                        for (int i = 0; i < PV; ++i) lds_st<C>(lds + v * vstride, (sl ^ S) ^ (i * EB), xc[i]);
                    }
                }
        }
    __syncthreads();
    int gi = 0;
    const int ob = uni(tk.st.op_begin), oe = uni(tk.st.op_end);
    if constexpr (VAR == 4) {
        if constexpr (NV == 1 && std::is_same<T, float>::value) run_reg_groups<CP>(lds, rprog, rfrag, rfa, rfl);
    } else {
    // lane-table words of op o (3-qubit MFMA ops), requested one op ahead
    constexpr bool PF = VAR >= 2 && std::is_same<T, float>::value;
    auto lane_words = [&](int o) {
        LaneWords w;
        w.on = false;
        if constexpr (PF) {
            if (o < oe && ltab) {
                const DESV_CONST DevOp *p = ops + o;
                const int lt = p->lt;
                if (p->kind == OPK_MFMA && p->k == 3 && lt >= 0)
                    w = load_lane_words<CP>(ltab + (size_t)lt * (kLaneTableLanes * kLaneTableWords / 4));
            }
        }
        return w;
    };
    LaneWords pf = lane_words(ob);
    for (int o = ob; o < oe; ++o) {
        const DevOp op = gload_op(ops + o);
        const LaneWords lw = pf;
        if constexpr (PF) pf = lane_words(o + 1);
        const bool is_grad = op.kind == OPK_GRAD || op.kind == OPK_GRADDIAG;
        if ((base & op.cglob) != op.cglob) {
            gi += is_grad;
            continue;
        }
        const DESV_GLOBAL C *M = mats + op.mat, *ML = matl + op.mat;
        switch (op.kind) {
            case OPK_DENSE1: op_dense_lo<T, NV, 1, CP>(lds, vstride, m, op, M, ML); break;
            case OPK_DENSE2: op_dense_lo<T, NV, 2, CP>(lds, vstride, m, op, M, ML); break;
            case OPK_DENSEK:
                if constexpr (VAR == 1) {
                    if (op.k == 3) op_densek<T, NV, 3, CP>(lds, vstride, m, op, M, ML);
                    else if (op.k == 4) op_densek<T, NV, 4, CP>(lds, vstride, m, op, M, ML);
                    else op_densek<T, NV, 5, CP>(lds, vstride, m, op, M, ML);
                } else if constexpr (VAR >= 2) {
                    op_densek<T, NV, 3, CP>(lds, vstride, m, op, M, ML);
                }
                break;
            case OPK_MFMA:
                if constexpr (VAR >= 2 && std::is_same<T, float>::value) {
                    if (op.k == 3)
                        op_mfma3<NV, CP>(lds, vstride, m, op, M, ML, lw);
                    else if (op.k == 4) op_mfma<NV, 4, CP>(lds, vstride, m, op, M, ML);
                    else if constexpr (VAR == 3) op_mfma<NV, 5, CP>(lds, vstride, m, op, M, ML);
                }
                break;
            case OPK_DIAG:
                if (op.k == 1) op_diag_lo<T, NV, 1, CP>(lds, vstride, m, op, M, ML, base);
                else op_diag<T, NV, CP>(lds, vstride, m, op, M, ML, base);
                break;
            case OPK_SCALAR: op_scalar<T, NV, CP>(lds, vstride, m, M, ML); break;
            case OPK_GRAD:
            case OPK_GRADDIAG:
                if constexpr (NV == 2) {
                    double part;
                    if (op.pos[4] == 1) {  // fused with the next op (a block containing the probe's bit)
                        const DevOp nx = gload_op(ops + o + 1);
                        part = op_grad_block<T, CP>(lds, vstride, m, op, M, nx, mats + nx.mat, matl + nx.mat);
                        ++o;
                        if constexpr (PF) pf = lane_words(o + 1);  // the block was op o + 1
                    } else {
                        part = op.kind == OPK_GRAD ? op_grad_dense<T>(lds, vstride, m, op, M)
                                                   : op_grad_diag<T>(lds, vstride, m, op, M);
                    }
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
    }
    // store (through the write-back bit permutation when `permuted`), LB
    // chunks of LDS reads before their stores
    if (lane_on)
        for (int jb = 0; jb < nchunks; jb += LB) {
            V x[LB][NV];
#pragma unroll
            for (int u = 0; u < LB; ++u)
                if (jb + u < nchunks) {
                    const int Q = qtab[jb + u];
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        C *xc = reinterpret_cast<C *>(&x[u][v]);
#pragma unroll
                        for (int i = 0; i < PV; ++i) xc[i] = lds_ld<C>(lds + v * vstride, ql[i] ^ Q);
                    }
                }
#pragma unroll
            for (int u = 0; u < LB; ++u)
                if (jb + u < nchunks) {
                    const uint64_t off = gl | jtab[jb + u];
#pragma unroll
                    for (int v = 0; v < NV; ++v) gstore((DESV_GLOBAL V *)(svp[v] + off), x[u][v]);
                }
        }
    if (n_grad > 0) {
        __syncthreads();
        int g2 = 0;
        for (int o = ob; o < oe; ++o) {
            const int kind = gload_op(ops + o).kind;
            if (kind == OPK_GRAD || kind == OPK_GRADDIAG) {
                if (threadIdx.x == 0) {
                    double sum = 0;
                    for (int w = 0; w < kThreads / 64; ++w) sum += gacc[g2 * (kThreads / 64) + w];
                    tk.partials[(int64_t)gload_op(ops + o).slot * tk.pstride + t] = sum;
                }
                ++g2;
            }
        }
    }
}

template <typename T>
hipError_t launch_stage(int nv, int variant, bool comp, const StageTask *tasks, const int32_t *wg_task,
                        const StageTask &single, int64_t grid, int max_m, hipStream_t s) {
    using C = typename VT<T>::c;
    // the tile(s), then the bookkeeping
    const size_t tiles = ((size_t)nv * ((size_t)1 << max_m) * sizeof(C) + 15) / 16 * 16;
    const size_t lds = tiles + kStageLdsExtra;
    const int xoff = (int)tiles;
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffffll) return hipErrorInvalidValue;
    if (variant < 0 || variant > 4) return hipErrorInvalidValue;
    // register groups: complex64 forward stages only
    if (variant == 4 && (!std::is_same<T, float>::value || nv != 1)) return hipErrorInvalidValue;
    // complex128 never uses MFMA blocks: its variants 2 and 3 would only
    // differ from 1 by omitting k = 4, 5 VALU blocks
    if (!std::is_same<T, float>::value && variant >= 2) variant = 1;
    // compensated matrices: complex64 only
    if (comp && !std::is_same<T, float>::value) return hipErrorInvalidValue;
    const dim3 g((unsigned)grid);
#define DESV_ST(NVV, VV) \
    hipLaunchKernelGGL((k_stage<T, NVV, VV, false>), g, dim3(kThreads), lds, s, tasks, wg_task, single, xoff)
#define DESV_STC(NVV, VV) \
    hipLaunchKernelGGL((k_stage<float, NVV, VV, true>), g, dim3(kThreads), lds, s, tasks, wg_task, single, xoff)
    if constexpr (std::is_same<T, float>::value) {
        if (variant == 4) {
            if (comp) DESV_STC(1, 4); else DESV_ST(1, 4);
            return hipGetLastError();
        }
        if (comp) {
            if (nv == 1) {
                if (variant == 0) DESV_STC(1, 0); else if (variant == 1) DESV_STC(1, 1);
                else if (variant == 2) DESV_STC(1, 2); else DESV_STC(1, 3);
            } else {
                if (variant == 0) DESV_STC(2, 0); else if (variant == 1) DESV_STC(2, 1);
                else if (variant == 2) DESV_STC(2, 2); else DESV_STC(2, 3);
            }
            return hipGetLastError();
        }
    }
    if (nv == 1) {
        if (variant == 0) DESV_ST(1, 0); else if (variant == 1) DESV_ST(1, 1);
        else if (variant == 2) DESV_ST(1, 2); else DESV_ST(1, 3);
    } else {
        if (variant == 0) DESV_ST(2, 0); else if (variant == 1) DESV_ST(2, 1);
        else if (variant == 2) DESV_ST(2, 2); else DESV_ST(2, 3);
    }
#undef DESV_ST
#undef DESV_STC
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

// lambda = H psi for one aligned tile of 2^tb amplitudes per workgroup
// (tb = log2 chunk: 11, fewer for small states).  Flip-mask groups arrive sorted by mask, so the groups sharing the
// mask's bits above the tile are consecutive: their gathers psi[i ^ x] all
// fall in one aligned input tile, loaded into LDS once per such run instead
// of once per group.  The sign (-1)^popc((i ^ x) & z) splits into the part
// above the tile (uniform, folded into the coefficient) and the part inside.
// FP64 accumulation; lambda (if requested), the energy partial <psi|lambda>
// and the psi copy (if requested) are formed at the end.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_matvec_tasks(const MatvecTask *__restrict__ tasks, int ntasks) {
    using C = typename VT<T>::c;
    constexpr int TB = std::is_same<T, float>::value ? kMatvecTileBitsC64 : kMatvecTileBitsC128;
    constexpr int E = (1 << TB) / kThreads;  // amplitudes per lane
    __shared__ double2 tile[1 << TB];  // converted once per load
    __shared__ double red[kThreads / 64];
    const MatvecTask tk = tasks[find_task(tasks, ntasks)];
    const int64_t w = (int64_t)blockIdx.x - tk.wg_begin;
    const int64_t tsize = tk.chunk;  // power of two <= 2^TB
    const uint64_t lomask = (uint64_t)tsize - 1;
    const uint64_t i0 = (uint64_t)w * (uint64_t)tsize;
    const C *psi = reinterpret_cast<const C *>(tk.psi);
    C *lam = reinterpret_cast<C *>(tk.lam);
    C *psi2 = reinterpret_cast<C *>(tk.psi2);
    const DESV_GLOBAL MatvecGroup *groups = gptr(reinterpret_cast<const MatvecGroup *>(tk.groups));
    const DESV_GLOBAL uint64_t *zt = gptr(tk.z);
    const DESV_GLOBAL double2 *ct = gptr(tk.cc);
    const int ng = uni(tk.ngroups);
    const int t = threadIdx.x;
    double2 acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = make_double2(0.0, 0.0);
    uint64_t cur = ~0ull;
    for (int g = 0; g < ng; ++g) {
        const uint64_t x = uni64(groups[g].x);
        const uint64_t xh = x & ~lomask;
        const uint32_t xl = (uint32_t)(x & lomask);
        if (xh != cur) {
            __syncthreads();  // previous tile fully read
            const C *src = psi + (i0 ^ xh);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int j = t + e * kThreads;
                if (j < tsize) tile[j] = to_d(src[j]);
            }
            __syncthreads();
            cur = xh;
        }
        double2 a[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int j = t + e * kThreads;
            a[e] = j < tsize ? tile[j ^ xl] : make_double2(0.0, 0.0);
        }
        const uint64_t hi = i0 ^ xh;  // bits of i ^ x above the tile
        const int qb = uni(groups[g].begin), qe = uni(groups[g].end);
        for (int q = qb; q < qe; ++q) {
            const uint64_t z = uni64(zt[q]);
            double2 c = gload(ct + q);
            if (__popcll(hi & z) & 1) c = make_double2(-c.x, -c.y);
            const uint32_t zl = (uint32_t)(z & lomask);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const uint32_t jl = (uint32_t)(t + e * kThreads) ^ xl;
                const double sg = (__popc(jl & zl) & 1) ? -1.0 : 1.0;
                const double2 ad = a[e];
                const double pr = c.x * ad.x - c.y * ad.y, pi = c.x * ad.y + c.y * ad.x;
                acc[e].x = __builtin_fma(sg, pr, acc[e].x);
                acc[e].y = __builtin_fma(sg, pi, acc[e].y);
            }
        }
    }
    // the energy partial and the psi copy only when asked for (uniform)
    const bool full = tk.epart != nullptr || psi2 != nullptr;
    double en = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = t + e * kThreads;
        if (j >= tsize) continue;
        if (lam) {
            C r;
            r.x = (T)acc[e].x;
            r.y = (T)acc[e].y;
            lam[i0 + j] = r;
        }
        if (full) {
            const C p = psi[i0 + j];
            const double2 pd = to_d(p);
            en += pd.x * acc[e].x + pd.y * acc[e].y;
            if (psi2) psi2[i0 + j] = p;
        }
    }
    if (tk.epart) {
        const double sum = block_sum(en, red);
        if (threadIdx.x == 0) tk.epart[w] = sum;
    }
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
    if (threadIdx.x == 0) {
        *tk.dst = s;
        if (tk.stamp) atomicMax(tk.stamp, (unsigned long long)wall_clock64());
    }
}

__global__ void __launch_bounds__(kThreads) k_mark(unsigned long long *stamps, int64_t count) {
    for (int64_t i = threadIdx.x; i < count; i += blockDim.x) stamps[1 + i] = 0ull;
    if (threadIdx.x == 0) stamps[0] = (unsigned long long)wall_clock64();
}

hipError_t launch_mark(unsigned long long *stamps, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_mark, dim3(1), dim3(kThreads), 0, s, stamps, count);
    return hipGetLastError();
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

// OOP = true reads `src` and writes `dst` (a second buffer of the same size):
// amplitudes whose controls are not satisfied are copied unchanged.  OOP =
// false updates `dst` (== src) in place and skips them.
template <typename T, int K, int VEC, bool DIAG, int GPT, bool NT, bool OOP>
__global__ void __launch_bounds__(kThreads) k_direct(const typename VT<T>::c *src, typename VT<T>::c *dst,
                                                    int64_t groups, DirectOp op,
                                                    const typename VT<T>::c *__restrict__ M, int xcd, int contig) {
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
    int64_t bid = blockIdx.x;
    if (xcd > 1) {
        // workgroups are dealt to the XCDs round-robin: give XCD x the x-th
        // contiguous slice of the grid's blocks (a bijection; a tail that
        // does not fill all XCDs keeps its blocks)
        const int64_t per = (int64_t)gridDim.x / xcd;
        if (bid < per * xcd) bid = (bid % xcd) * per + bid / xcd;
    }
    const int64_t t0 = bid * blockDim.x + threadIdx.x;
    uint64_t b[GPT];
    bool on[GPT], ctl[GPT];
    C a[GPT][D][VEC];
#pragma unroll
    for (int u = 0; u < GPT; ++u) {
        // contig: wave w owns the GPT consecutive 64-vector runs w GPT, ..., w GPT + GPT - 1
        const int64_t g0 = (contig ? ((t0 >> 6) * GPT + u) * 64 + (t0 & 63) : t0 + u * nthr) * VEC;
        uint64_t x = (uint64_t)g0;
#pragma unroll
        for (int j = 0; j < K; ++j) x = ((x >> srt[j]) << (srt[j] + 1)) | (x & ((1ull << srt[j]) - 1));
        b[u] = x;
        ctl[u] = (x & cm) == cm;
        on[u] = g0 < groups && (OOP || ctl[u]);
        if (!on[u]) continue;
#pragma unroll
        for (int s = 0; s < D; ++s) {
            if constexpr (VEC == 2) {
                const float4 v = ld<float4, NT>(reinterpret_cast<const float4 *>(&src[x | off[s]]));
                a[u][s][0] = make_float2(v.x, v.y);
                a[u][s][1] = make_float2(v.z, v.w);
            } else {
                a[u][s][0] = ld<C, NT>(&src[x | off[s]]);
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
                if (OOP && !ctl[u]) {
                    r[q][w] = a[u][q][w];
                } else if constexpr (DIAG) {
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
                st<float4, NT>(reinterpret_cast<float4 *>(&dst[b[u] | off[s]]),
                               make_float4(r[s][0].x, r[s][0].y, r[s][1].x, r[s][1].y));
            } else {
                st<C, NT>(&dst[b[u] | off[s]], r[s][0]);
            }
        }
    }
}

namespace {
// Launch shape of the direct kernel: group vectors per thread (GPT), whether
// a wave's GPT vectors are consecutive 64-vector runs (contig) or strided by
// the grid, non-temporal access (nt), and XCD-sliced workgroup order.
// DESV_DIRECT_GPT / _CONTIG / _NT / _XCD override the per-op choice.
int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}
int direct_xcd() {
    // MI355X deals workgroups to its 8 XCDs round-robin; slicing the grid per
    // XCD measured 5.38 -> 5.52 TB/s on 30-qubit 2-qubit-gate passes.
    static int v = env_int("DESV_DIRECT_XCD", 8);
    return v;
}
struct DirectShape {
    int gpt;
    bool nt, contig;
};
// By the lowest target bit lo, ms per pass of a 2-qubit brickwork gate
// (strided GPT 1 / contig GPT 4 / contig GPT 4 + nt):
//   complex64, 2 amplitudes per 16-byte access, 30 qubits:
//     lo = 1: 3.74 / 3.68 / 11.75    lo = 2: 3.43 / 3.23 / 8.44
//     lo = 4..8: 2.93 / 2.98 / 2.96  lo >= 9: 2.67..3.60, nt + contig 2-6 % faster
//   complex128, 1 amplitude per access, 29 qubits:
//     lo = 0: 4.01 / 3.76 / 10.66    lo = 1: 3.71 / 3.37 / 7.14
//     lo = 3..9: 2.95..2.97 / 2.90..3.00 / 2.76..2.81; lo >= 10: nt + contig mostly best.
// Non-temporal stores need whole 128-byte lines per store instruction.
DirectShape direct_shape(bool dbl, bool vec16, int lo) {
    static const int g = env_int("DESV_DIRECT_GPT", 0), c = env_int("DESV_DIRECT_CONTIG", -1),
                     t = env_int("DESV_DIRECT_NT", -1);
    DirectShape d{1, false, false};
    if (dbl) d = lo >= 3 ? DirectShape{4, true, true} : DirectShape{4, false, true};
    else if (vec16 && lo >= 9) d = {4, true, true};
    else if (vec16 && lo <= 2) d = {4, false, true};
    if (g == 1 || g == 2 || g == 4) d.gpt = g;
    if (c >= 0) d.contig = c != 0;
    if (t >= 0) d.nt = t != 0;
    return d;
}
// The lane-exchange kernel for targets in bits 1..6: used where the strided
// 16-byte kernel is slow (lowest target bit 1 or 2); DESV_DIRECT_LANE = 0 / 1
// turns it off / on for every eligible op.
bool direct_lane(int lo) {
    static const int v = env_int("DESV_DIRECT_LANE", -1);
    return v < 0 ? lo <= 2 : v != 0;
}
}  // namespace

// Complex64 gates whose targets or controls include index bit 0, still with
// 16-byte accesses: a float4 holds the amplitude pair (x, x | 1), so bit 0
// stays inside the lane.  J0 >= 0: bit 0 is target J0 of the op (matrix bit
// J0); J0 < 0: bit 0 is a control.  The other NH targets are "outer" bits:
// thread u owns the 2^NH float4 pairs at x | off(c), c < 2^NH, and amplitude
// (c, h) = amp[x | off(c) | h].  With bit 0 a control only the h = 1 half is
// transformed; the h = 0 half is written back as read.
template <int K, int J0, bool DIAG, bool OOP>
__global__ void __launch_bounds__(kThreads) k_direct_b0(const float2 *src, float2 *dst, int64_t units, DirectOp op,
                                                       const float2 *__restrict__ M, int xcd) {
    constexpr int D = 1 << K;
    constexpr int NH = J0 < 0 ? K : K - 1;
    constexpr int NC = 1 << NH;
    // matrix bit of outer bit i
    auto mbit = [](int i) constexpr { return (J0 >= 0 && i >= J0) ? i + 1 : i; };
    uint64_t off[NC];
    int lo = 64, hi = 0;  // outer positions, ascending (NH <= 2)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        uint64_t o = 0;
#pragma unroll
        for (int i = 0; i < NH; ++i)
            if (c >> i & 1) o |= 1ull << op.pos[mbit(i)];
        off[c] = o;
    }
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        lo = min(lo, op.pos[mbit(i)]);
        hi = max(hi, op.pos[mbit(i)]);
    }
    float2 mm[DIAG ? D : D * D];
#pragma unroll
    for (int i = 0; i < (DIAG ? D : D * D); ++i) mm[i] = M[i];
    const uint64_t cm = op.cmask;
    int64_t bid = blockIdx.x;
    if (xcd > 1) {
        const int64_t per = (int64_t)gridDim.x / xcd;
        if (bid < per * xcd) bid = (bid % xcd) * per + bid / xcd;
    }
    const int64_t u = bid * blockDim.x + threadIdx.x;
    if (u >= units) return;
    // pair index y (bit b of y = index bit b + 1) with zeros at the outer bits
    uint64_t y = (uint64_t)u;
    if constexpr (NH >= 1) y = ((y >> (lo - 1)) << lo) | (y & ((1ull << (lo - 1)) - 1));
    if constexpr (NH == 2) y = ((y >> (hi - 1)) << hi) | (y & ((1ull << (hi - 1)) - 1));
    const uint64_t x = y << 1;
    const bool act = J0 >= 0 ? (x & cm) == cm : ((x | 1) & cm) == cm;
    if (!OOP && !act) return;
    float2 a[NC][2];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const float4 v = ld<float4, false>(reinterpret_cast<const float4 *>(&src[x | off[c]]));
        a[c][0] = make_float2(v.x, v.y);
        a[c][1] = make_float2(v.z, v.w);
    }
    float2 r[NC][2];
#pragma unroll
    for (int c = 0; c < NC; ++c) r[c][0] = a[c][0], r[c][1] = a[c][1];
    if (act) {
        if constexpr (J0 >= 0) {
            float2 in[D];  // in[matrix index]
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    int q = h << J0;
#pragma unroll
                    for (int i = 0; i < NH; ++i) q |= (c >> i & 1) << mbit(i);
                    in[q] = a[c][h];
                }
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    int q = h << J0;
#pragma unroll
                    for (int i = 0; i < NH; ++i) q |= (c >> i & 1) << mbit(i);
                    if constexpr (DIAG) {
                        r[c][h] = cmul(mm[q], in[q]);
                    } else {
                        float2 acc = cmul(mm[q * D], in[0]);
#pragma unroll
                        for (int k = 1; k < D; ++k) acc = cfma(mm[q * D + k], in[k], acc);
                        r[c][h] = acc;
                    }
                }
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) {  // matrix index of (c, 1) is c
                if constexpr (DIAG) {
                    r[c][1] = cmul(mm[c], a[c][1]);
                } else {
                    float2 acc = cmul(mm[c * D], a[0][1]);
#pragma unroll
                    for (int k = 1; k < D; ++k) acc = cfma(mm[c * D + k], a[k][1], acc);
                    r[c][1] = acc;
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c)
        st<float4, false>(reinterpret_cast<float4 *>(&dst[x | off[c]]),
                          make_float4(r[c][0].x, r[c][0].y, r[c][1].x, r[c][1].y));
}

template <int K, int J0, bool DIAG>
hipError_t launch_direct_b0(const float2 *src, float2 *dst, int n, const DirectOp &op, const float2 *m, hipStream_t s) {
    constexpr int NH = J0 < 0 ? K : K - 1;
    const int64_t units = 1ll << (n - 1 - NH);
    const int64_t grid = std::max<int64_t>(1, (units + kThreads - 1) / kThreads);
    if (grid > 0x7fffffffll) return hipErrorInvalidValue;
    if (src != dst)
        hipLaunchKernelGGL((k_direct_b0<K, J0, DIAG, true>), dim3((unsigned)grid), dim3(kThreads), 0, s, src, dst,
                           units, op, m, direct_xcd());
    else
        hipLaunchKernelGGL((k_direct_b0<K, J0, DIAG, false>), dim3((unsigned)grid), dim3(kThreads), 0, s, src, dst,
                           units, op, m, direct_xcd());
    return hipGetLastError();
}

template <int K, bool DIAG>
hipError_t dispatch_direct_b0(const float2 *src, float2 *dst, int n, const DirectOp &op, const float2 *m,
                              hipStream_t s) {
    if (op.pos[0] == 0) return launch_direct_b0<K, 0, DIAG>(src, dst, n, op, m, s);
    if constexpr (K == 2)
        if (op.pos[1] == 0) return launch_direct_b0<K, 1, DIAG>(src, dst, n, op, m, s);
    return launch_direct_b0<K, -1, DIAG>(src, dst, n, op, m, s);
}

// Complex64 gates whose targets all lie in index bits 1..6 (no target at
// bit 0): every access instruction is lane-contiguous (wave w, step s, lane l
// owns float4 f = 256 w + 64 s + l, the amplitude pair (2f, 2f + 1)), so the
// members of a group sit in lanes that differ at lane bits pos - 1 and are
// exchanged by xor shuffles.  A lane's member index m depends on the lane
// only, so its matrix row M[m][m ^ d] is loaded once.  Controls never touch
// target bits, so a group shares its control outcome (per half: bit 0 may be
// a control).  nvec4 must be a multiple of 256 (whole waves in range).
template <int K, bool DIAG, bool OOP>
__global__ void __launch_bounds__(kThreads) k_direct_lane(const float2 *src, float2 *dst, int64_t nvec4, DirectOp op,
                                                         const float2 *__restrict__ M, int xcd) {
    constexpr int D = 1 << K, S = 4;
    const int lane = threadIdx.x & 63;
    int m = 0, dep[D];
#pragma unroll
    for (int j = 0; j < K; ++j) m |= ((lane >> (op.pos[j] - 1)) & 1) << j;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        int v = 0;
#pragma unroll
        for (int j = 0; j < K; ++j) v |= ((d >> j) & 1) << (op.pos[j] - 1);
        dep[d] = v;
    }
    float2 c[DIAG ? 1 : D];
    if constexpr (DIAG) {
        c[0] = M[m];
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) c[d] = M[m * D + (m ^ d)];
    }
    int64_t bid = blockIdx.x;
    if (xcd > 1) {
        const int64_t per = (int64_t)gridDim.x / xcd;
        if (bid < per * xcd) bid = (bid % xcd) * per + bid / xcd;
    }
    const int64_t f0 = ((bid * blockDim.x + threadIdx.x) >> 6) * (S * 64);
    if (f0 >= nvec4) return;  // the whole wave
    const uint64_t cm = op.cmask;
    float4 a[S];
#pragma unroll
    for (int s = 0; s < S; ++s) a[s] = ld<float4, false>(reinterpret_cast<const float4 *>(src) + f0 + s * 64 + lane);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint64_t e0 = (uint64_t)(f0 + s * 64 + lane) << 1;
        const bool c0 = (e0 & cm) == cm, c1 = ((e0 | 1) & cm) == cm;
        float2 lo = make_float2(a[s].x, a[s].y), hi = make_float2(a[s].z, a[s].w);
        float2 r0 = cmul(c[0], lo), r1 = cmul(c[0], hi);
        if constexpr (!DIAG) {
#pragma unroll
            for (int d = 1; d < D; ++d) {
                const float2 plo = make_float2(__shfl_xor(a[s].x, dep[d]), __shfl_xor(a[s].y, dep[d]));
                const float2 phi = make_float2(__shfl_xor(a[s].z, dep[d]), __shfl_xor(a[s].w, dep[d]));
                r0 = cfma(c[d], plo, r0);
                r1 = cfma(c[d], phi, r1);
            }
        }
        if (!c0) r0 = lo;
        if (!c1) r1 = hi;
        if (OOP || c0 || c1)
            st<float4, false>(reinterpret_cast<float4 *>(dst) + f0 + s * 64 + lane, make_float4(r0.x, r0.y, r1.x, r1.y));
    }
}

template <int K, bool DIAG>
hipError_t launch_direct_lane(const float2 *src, float2 *dst, int n, const DirectOp &op, const float2 *m,
                              hipStream_t s) {
    const int64_t nvec4 = 1ll << (n - 1);
    const int64_t grid = std::max<int64_t>(1, nvec4 / 4 / kThreads + (nvec4 / 4 % kThreads != 0));
    if (grid > 0x7fffffffll) return hipErrorInvalidValue;
    if (src != dst)
        hipLaunchKernelGGL((k_direct_lane<K, DIAG, true>), dim3((unsigned)grid), dim3(kThreads), 0, s, src, dst,
                           nvec4, op, m, direct_xcd());
    else
        hipLaunchKernelGGL((k_direct_lane<K, DIAG, false>), dim3((unsigned)grid), dim3(kThreads), 0, s, src, dst,
                           nvec4, op, m, direct_xcd());
    return hipGetLastError();
}

template <typename T, int K, int VEC, bool DIAG>
hipError_t launch_direct_k(const typename VT<T>::c *src, typename VT<T>::c *dst, int64_t groups, const DirectOp &op,
                           const typename VT<T>::c *m, hipStream_t s) {
    int lo = 64;
    for (int j = 0; j < K; ++j) lo = std::min(lo, (int)op.pos[j]);
    const DirectShape shape = direct_shape(!std::is_same<T, float>::value, VEC == 2, lo);
    const int gpt = shape.gpt;
    const bool nt = shape.nt;
    const int contig = shape.contig;
    const int xcd = direct_xcd();
    const int64_t vecs = (groups + VEC - 1) / VEC;
    const int64_t threads = (vecs + gpt - 1) / gpt;
    const int64_t grid = std::max<int64_t>(1, (threads + kThreads - 1) / kThreads);
    if (grid > 0x7fffffffll) return hipErrorInvalidValue;
#define DESV_DK(G, N, O)                                                                               \
    hipLaunchKernelGGL((k_direct<T, K, VEC, DIAG, G, N, O>), dim3((unsigned)grid), dim3(kThreads), 0, s, src, dst, \
                       groups, op, m, xcd, contig)
    if (src != dst) {
        if (nt) {
            if (gpt == 1) DESV_DK(1, true, true); else if (gpt == 2) DESV_DK(2, true, true); else DESV_DK(4, true, true);
        } else {
            if (gpt == 1) DESV_DK(1, false, true); else if (gpt == 2) DESV_DK(2, false, true); else DESV_DK(4, false, true);
        }
    } else if (nt) {
        if (gpt == 1) DESV_DK(1, true, false); else if (gpt == 2) DESV_DK(2, true, false); else DESV_DK(4, true, false);
    } else {
        if (gpt == 1) DESV_DK(1, false, false); else if (gpt == 2) DESV_DK(2, false, false); else DESV_DK(4, false, false);
    }
#undef DESV_DK
    return hipGetLastError();
}

template <typename T>
hipError_t launch_direct(const void *in, void *out, int n, const DirectOp &op, const void *M, hipStream_t s) {
    using C = typename VT<T>::c;
    const int K = op.k;
    const int64_t groups = 1ll << (n - K);
    bool bit0 = (op.cmask & 1) != 0;
    for (int j = 0; j < K; ++j) bit0 |= op.pos[j] == 0;
    const bool vec = std::is_same<T, float>::value && !bit0 && groups >= 2;
    const C *p = reinterpret_cast<const C *>(in);
    C *q = reinterpret_cast<C *>(out);
    const C *m = reinterpret_cast<const C *>(M);
    if constexpr (std::is_same<T, float>::value) {
        if (bit0) {
            if (K == 1) return op.diag ? dispatch_direct_b0<1, true>(p, q, n, op, m, s)
                                       : dispatch_direct_b0<1, false>(p, q, n, op, m, s);
            return op.diag ? dispatch_direct_b0<2, true>(p, q, n, op, m, s)
                           : dispatch_direct_b0<2, false>(p, q, n, op, m, s);
        }
        int lo = 64, hi = 0;
        for (int j = 0; j < K; ++j) lo = std::min(lo, (int)op.pos[j]), hi = std::max(hi, (int)op.pos[j]);
        if (lo >= 1 && hi <= 6 && n >= 9 && direct_lane(lo)) {
            if (K == 1) return op.diag ? launch_direct_lane<1, true>(p, q, n, op, m, s)
                                       : launch_direct_lane<1, false>(p, q, n, op, m, s);
            return op.diag ? launch_direct_lane<2, true>(p, q, n, op, m, s)
                           : launch_direct_lane<2, false>(p, q, n, op, m, s);
        }
        if (K == 1) {
            if (op.diag) return vec ? launch_direct_k<T, 1, 2, true>(p, q, groups, op, m, s)
                                    : launch_direct_k<T, 1, 1, true>(p, q, groups, op, m, s);
            return vec ? launch_direct_k<T, 1, 2, false>(p, q, groups, op, m, s)
                       : launch_direct_k<T, 1, 1, false>(p, q, groups, op, m, s);
        }
        if (op.diag) return vec ? launch_direct_k<T, 2, 2, true>(p, q, groups, op, m, s)
                                : launch_direct_k<T, 2, 1, true>(p, q, groups, op, m, s);
        return vec ? launch_direct_k<T, 2, 2, false>(p, q, groups, op, m, s)
                   : launch_direct_k<T, 2, 1, false>(p, q, groups, op, m, s);
    } else {
        if (K == 1) return op.diag ? launch_direct_k<T, 1, 1, true>(p, q, groups, op, m, s)
                                   : launch_direct_k<T, 1, 1, false>(p, q, groups, op, m, s);
        return op.diag ? launch_direct_k<T, 2, 1, true>(p, q, groups, op, m, s)
                       : launch_direct_k<T, 2, 1, false>(p, q, groups, op, m, s);
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
    const void *fns[24] = {
        (const void *)k_stage<float, 1, 0, false>,  (const void *)k_stage<float, 2, 0, false>,
        (const void *)k_stage<double, 1, 0, false>, (const void *)k_stage<double, 2, 0, false>,
        (const void *)k_stage<float, 1, 1, false>,  (const void *)k_stage<float, 2, 1, false>,
        (const void *)k_stage<double, 1, 1, false>, (const void *)k_stage<double, 2, 1, false>,
        (const void *)k_stage<float, 1, 2, false>,  (const void *)k_stage<float, 2, 2, false>,
        (const void *)k_stage<double, 1, 2, false>, (const void *)k_stage<double, 2, 2, false>,
        (const void *)k_stage<float, 1, 3, false>,  (const void *)k_stage<float, 2, 3, false>,
        (const void *)k_stage<double, 1, 3, false>, (const void *)k_stage<double, 2, 3, false>,
        (const void *)k_stage<float, 1, 0, true>,   (const void *)k_stage<float, 1, 1, true>,
        (const void *)k_stage<float, 1, 2, true>,   (const void *)k_stage<float, 1, 3, true>,
        (const void *)k_stage<float, 2, 0, true>,   (const void *)k_stage<float, 2, 1, true>,
        (const void *)k_stage<float, 2, 2, true>,   (const void *)k_stage<float, 2, 3, true>};
    // the stage kernels' bookkeeping sits behind the tiles in dynamic LDS
    // (they have no static LDS): the tiles get what is left of it
    size_t static_lds = 0;
    for (const void *f : fns) {
        hipFuncAttributes a;
        if (hipFuncGetAttributes(&a, f) == hipSuccess) static_lds = std::max(static_lds, (size_t)a.sharedSizeBytes);
        else (void)hipGetLastError();
    }
    static_lds = (static_lds + kStageLdsExtra + 255) & ~(size_t)255;
    lim -= static_lds;
    for (const void *f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(lim + kStageLdsExtra)) !=
            hipSuccess) {
            (void)hipGetLastError();
            lim = 64 * 1024 - static_lds;
        }
    cached[device] = lim;
    return lim;
}

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------
// dst <- src, 16 bytes per lane, one access per lane (the streaming copy
// that reaches the highest HBM rate on MI355X; hipMemcpy device-to-device is
// slower).  bytes must be a multiple of 16.
__global__ void __launch_bounds__(kThreads) k_copy16(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     int64_t n16, int xcd) {
    int64_t bid = blockIdx.x;
    if (xcd > 1) {
        const int64_t per = (int64_t)gridDim.x / xcd;
        if (bid < per * xcd) bid = (bid % xcd) * per + bid / xcd;
    }
    const int64_t i = bid * blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}
hipError_t launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s) {
    const int64_t n16 = (int64_t)(bytes / 16);
    const int64_t grid = (n16 + kThreads - 1) / kThreads;
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffffll) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)grid), dim3(kThreads), 0, s, reinterpret_cast<const float4 *>(src),
                       reinterpret_cast<float4 *>(dst), n16, direct_xcd());
    return hipGetLastError();
}

// dst[i] = src[i] with a precision change (complex64 <-> complex128).
template <typename TD, typename TS>
__global__ void k_convert(typename VT<TD>::c *__restrict__ dst, const typename VT<TS>::c *__restrict__ src,
                          int64_t dim) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        const typename VT<TS>::c a = src[i];
        typename VT<TD>::c b;
        b.x = (TD)a.x;
        b.y = (TD)a.y;
        dst[i] = b;
    }
}
template <typename TD, typename TS>
hipError_t launch_convert(void *dst, const void *src, int64_t dim, hipStream_t s) {
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((dim + kThreads - 1) / kThreads, 1 << 20));
    hipLaunchKernelGGL((k_convert<TD, TS>), dim3((unsigned)grid), dim3(kThreads), 0, s,
                       reinterpret_cast<typename VT<TD>::c *>(dst), reinterpret_cast<const typename VT<TS>::c *>(src),
                       dim);
    return hipGetLastError();
}
template hipError_t launch_convert<float, double>(void *, const void *, int64_t, hipStream_t);
template hipError_t launch_convert<double, float>(void *, const void *, int64_t, hipStream_t);

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

// ---------------------------------------------------------------------------
// Shot sampling: the probability CDF of the state in three deterministic
// passes (per-chunk totals, one scan of the chunk totals, per-chunk scans
// with offset and normalisation), then one binary search per shot.
// ---------------------------------------------------------------------------
constexpr int kScanChunk = 4096;  // amplitudes per workgroup, 16 per lane

__device__ inline double block_inclusive_scan(double v, double *ts) {
    ts[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const double t = (int)threadIdx.x >= off ? ts[threadIdx.x - off] : 0.0;
        __syncthreads();
        ts[threadIdx.x] += t;
        __syncthreads();
    }
    return ts[threadIdx.x];
}

// First CDF pass: blk[chunk] = sum of |a|^2 over each kScanChunk-amplitude
// chunk, one workgroup per chunk, 16-byte coalesced loads (a float4 is two
// complex64 amplitudes or one complex128 amplitude) and no LDS staging.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_cdf_totals(const typename VT<T>::c *sv, int64_t dim, double *blk) {
    using C = typename VT<T>::c;
    constexpr int APU = 16 / (int)sizeof(C);  // amplitudes per 16-byte unit
    constexpr int per = kScanChunk / APU / kThreads;
    __shared__ double red[kThreads / 64];
    const int64_t c0 = (int64_t)blockIdx.x * kScanChunk;
    double acc = 0;
    if (c0 + kScanChunk <= dim) {
        const float4 *p = reinterpret_cast<const float4 *>(sv + c0);
        float4 v[per];
#pragma unroll
        for (int k = 0; k < per; ++k) v[k] = p[k * kThreads + threadIdx.x];
#pragma unroll
        for (int k = 0; k < per; ++k) {
            if constexpr (APU == 2) {
                const double2 a = to_d(make_float2(v[k].x, v[k].y)), b = to_d(make_float2(v[k].z, v[k].w));
                acc += a.x * a.x + a.y * a.y;
                acc += b.x * b.x + b.y * b.y;
            } else {
                double2 a;
                __builtin_memcpy(&a, &v[k], sizeof(a));
                acc += a.x * a.x + a.y * a.y;
            }
        }
    } else {
        for (int64_t e = c0 + threadIdx.x; e < dim; e += kThreads) {
            const double2 a = to_d(sv[e]);
            acc += a.x * a.x + a.y * a.y;
        }
    }
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) blk[blockIdx.x] = t;
}

// The CDF is built from fixed-point weights so that every association of
// the sums gives the same prefix: q_i = round(|a_i|^2 2^62 / total) as a
// 64-bit integer (total from the FP64 pass above), chunk totals and prefixes
// summed in integers, cdf_i = P_i / P_total in FP64.  Integer prefixes make
// the CDF nondecreasing across chunk boundaries and give a zero-probability
// amplitude exactly zero width, whatever the chunk and scan layout; the
// quantisation (2^-62 of the total per amplitude) is far below FP64's
// resolution of the prefix.
__device__ inline double cdf_scale(const double *blk, int64_t nb) {
    const double t = blk[nb];
    return (t > 0.0 && t <= 1.0e300) ? 4611686018427387904.0 / t : 0.0;  // 2^62 / total
}
template <typename C> __device__ inline uint64_t cdf_weight(C a, double scale) {
    const double2 d = to_d(a);
    return (uint64_t)((d.x * d.x + d.y * d.y) * scale + 0.5);
}

__device__ inline uint64_t block_sum_u64(uint64_t v, uint64_t *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    uint64_t s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    return s;
}

__device__ inline uint64_t block_inclusive_scan_u64(uint64_t v, uint64_t *ts) {
    ts[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const uint64_t t = (int)threadIdx.x >= off ? ts[threadIdx.x - off] : 0ull;
        __syncthreads();
        ts[threadIdx.x] += t;
        __syncthreads();
    }
    return ts[threadIdx.x];
}

// qblk[chunk] = sum of the chunk's fixed-point weights (16-byte coalesced
// loads, as in k_cdf_totals).
template <typename T>
__global__ void __launch_bounds__(kThreads) k_cdf_qtotals(const typename VT<T>::c *sv, int64_t dim, const double *blk,
                                                          int64_t nb, uint64_t *qblk) {
    using C = typename VT<T>::c;
    constexpr int APU = 16 / (int)sizeof(C);
    constexpr int per = kScanChunk / APU / kThreads;
    __shared__ uint64_t red[kThreads / 64];
    const double scale = cdf_scale(blk, nb);
    const int64_t c0 = (int64_t)blockIdx.x * kScanChunk;
    uint64_t acc = 0;
    if (c0 + kScanChunk <= dim) {
        const float4 *p = reinterpret_cast<const float4 *>(sv + c0);
        float4 v[per];
#pragma unroll
        for (int k = 0; k < per; ++k) v[k] = p[k * kThreads + threadIdx.x];
#pragma unroll
        for (int k = 0; k < per; ++k) {
            C a[APU];
            __builtin_memcpy(a, &v[k], sizeof(a));
#pragma unroll
            for (int j = 0; j < APU; ++j) acc += cdf_weight(a[j], scale);
        }
    } else {
        for (int64_t e = c0 + threadIdx.x; e < dim; e += kThreads) acc += cdf_weight(sv[e], scale);
    }
    const uint64_t t = block_sum_u64(acc, red);
    if (threadIdx.x == 0) qblk[blockIdx.x] = t;
}

// qblk[0..nb) <- exclusive prefix sums; qblk[nb] <- total.  One workgroup.
__global__ void __launch_bounds__(kThreads) k_scan_qblocks(uint64_t *qblk, int64_t nb) {
    __shared__ uint64_t ts[kThreads];
    const int64_t seg = (nb + kThreads - 1) / kThreads;
    const int64_t s0 = (int64_t)threadIdx.x * seg, s1 = s0 + seg < nb ? s0 + seg : nb;
    uint64_t run = 0;
    for (int64_t i = s0; i < s1; ++i) run += qblk[i];
    const uint64_t incl = block_inclusive_scan_u64(run, ts);
    uint64_t acc = incl - run;
    for (int64_t i = s0; i < s1; ++i) {
        const uint64_t x = qblk[i];
        qblk[i] = acc;
        acc += x;
    }
    if (threadIdx.x == blockDim.x - 1) qblk[nb] = incl;
}

// Final CDF pass, one 4096-amplitude chunk per workgroup: lane t sums the
// weights of its 16 consecutive amplitudes, a workgroup scan gives its
// prefix, and cdf = (prefix + chunk offset) / total.  Global loads and stores
// are coalesced through LDS (element e at e + e / 16: lane t's 16 elements
// then fall in distinct banks).
template <typename T>
__global__ void __launch_bounds__(kThreads) k_cdf_pass(const typename VT<T>::c *sv, int64_t dim, double *cdf,
                                                       const double *blk, const uint64_t *qblk, int64_t nb) {
    constexpr int per = kScanChunk / kThreads;
    __shared__ uint64_t ts[kThreads];
    __shared__ uint64_t buf[kScanChunk + kScanChunk / per];
    auto at = [](int e) { return e + e / per; };
    const double scale = cdf_scale(blk, nb);
    const int64_t c0 = (int64_t)blockIdx.x * kScanChunk;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int e = k * kThreads + t;
        buf[at(e)] = c0 + e < dim ? cdf_weight(sv[c0 + e], scale) : 0ull;
    }
    __syncthreads();
    uint64_t v[per], run = 0;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        run += buf[at(t * per + j)];
        v[j] = run;
    }
    const uint64_t incl = block_inclusive_scan_u64(run, ts);
    const uint64_t base = incl - run + qblk[blockIdx.x];
    const double total = (double)qblk[nb];
    __syncthreads();  // every lane has read its elements
    double *out = reinterpret_cast<double *>(buf);
#pragma unroll
    for (int j = 0; j < per; ++j) out[at(t * per + j)] = (double)(base + v[j]) / total;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int e = k * kThreads + t;
        if (c0 + e < dim) cdf[c0 + e] = out[at(e)];
    }
}

// blk[0..nb) <- exclusive prefix sums; blk[nb] <- total.  One workgroup.
__global__ void __launch_bounds__(kThreads) k_scan_blocks(double *blk, int64_t nb) {
    __shared__ double ts[kThreads];
    const int64_t seg = (nb + kThreads - 1) / kThreads;
    const int64_t s0 = (int64_t)threadIdx.x * seg, s1 = s0 + seg < nb ? s0 + seg : nb;
    double run = 0;
    for (int64_t i = s0; i < s1; ++i) run += blk[i];
    const double incl = block_inclusive_scan(run, ts);
    double acc = incl - run;
    for (int64_t i = s0; i < s1; ++i) {
        const double x = blk[i];
        blk[i] = acc;
        acc += x;
    }
    if (threadIdx.x == blockDim.x - 1) blk[nb] = incl;
}

// out[s] = number of cdf entries <= u[s] (the first index whose cumulative
// probability exceeds u[s]), clamped to dim - 1.
__global__ void k_sample_search(const double *cdf, int64_t dim, const double *u, int64_t shots, int64_t *out) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < shots; s += (int64_t)gridDim.x * blockDim.x) {
        const double x = u[s];
        int64_t lo = 0, hi = dim;
        while (lo < hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            if (cdf[mid] <= x) lo = mid + 1;
            else hi = mid;
        }
        out[s] = lo < dim ? lo : dim - 1;
    }
}

template <typename T>
hipError_t launch_cdf(const void *sv, int64_t dim, double *cdf, double *blk, hipStream_t s) {
    const int64_t nb = (dim + kScanChunk - 1) / kScanChunk;
    if (nb > 0x7fffffffll) return hipErrorInvalidValue;
    using C = typename VT<T>::c;
    uint64_t *qblk = reinterpret_cast<uint64_t *>(blk + nb + 1);
    hipLaunchKernelGGL((k_cdf_totals<T>), dim3((unsigned)nb), dim3(kThreads), 0, s, (const C *)sv, dim, blk);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kThreads), 0, s, blk, nb);
    hipLaunchKernelGGL((k_cdf_qtotals<T>), dim3((unsigned)nb), dim3(kThreads), 0, s, (const C *)sv, dim,
                       (const double *)blk, nb, qblk);
    hipLaunchKernelGGL(k_scan_qblocks, dim3(1), dim3(kThreads), 0, s, qblk, nb);
    hipLaunchKernelGGL((k_cdf_pass<T>), dim3((unsigned)nb), dim3(kThreads), 0, s, (const C *)sv, dim, cdf,
                       (const double *)blk, (const uint64_t *)qblk, nb);
    return hipGetLastError();
}
int64_t cdf_blocks(int64_t dim) { return (dim + kScanChunk - 1) / kScanChunk; }
hipError_t launch_sample_search(const double *cdf, int64_t dim, const double *u, int64_t shots, int64_t *out,
                                hipStream_t s) {
    if (shots <= 0) return hipSuccess;
    const int64_t grid = std::min<int64_t>((shots + kThreads - 1) / kThreads, 1 << 16);
    hipLaunchKernelGGL(k_sample_search, dim3((unsigned)grid), dim3(kThreads), 0, s, cdf, dim, u, shots, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Two-level shot sampling (no CDF array): one pass gives every chunk's FP64
// total and its fixed-point total with the fixed scale 2^62 (|a|^2 <= 1 per
// amplitude, so a state of total probability below 4 cannot overflow the
// 64-bit sums); a scan of the chunk totals; each shot's target T = floor(u
// P_total) picks the first chunk whose inclusive prefix exceeds T; shots are
// bucketed by chunk (counting sort), and one workgroup per chunk that holds
// shots rebuilds the chunk's integer prefix in LDS and binary-searches every
// shot's first index with prefix > T.  The state is read twice at most, the
// weights and prefixes are the integer ones of the CDF path, so an amplitude
// of probability zero has zero width.
// ---------------------------------------------------------------------------
constexpr double kSampleScale = 4611686018427387904.0;  // 2^62

template <typename T>
__global__ void __launch_bounds__(kThreads) k_sample_totals(const typename VT<T>::c *sv, int64_t dim, double *blk,
                                                            uint64_t *qblk) {
    using C = typename VT<T>::c;
    constexpr int APU = 16 / (int)sizeof(C);
    constexpr int per = kScanChunk / APU / kThreads;
    __shared__ double red[kThreads / 64];
    __shared__ uint64_t qred[kThreads / 64];
    const int64_t c0 = (int64_t)blockIdx.x * kScanChunk;
    double acc = 0;
    uint64_t qacc = 0;
    auto add = [&](C a) {
        const double2 d = to_d(a);
        const double p = d.x * d.x + d.y * d.y;
        acc += p;
        qacc += (uint64_t)(p * kSampleScale + 0.5);
    };
    if (c0 + kScanChunk <= dim) {
        const float4 *p = reinterpret_cast<const float4 *>(sv + c0);
        float4 v[per];
#pragma unroll
        for (int k = 0; k < per; ++k) v[k] = p[k * kThreads + threadIdx.x];
#pragma unroll
        for (int k = 0; k < per; ++k) {
            C a[APU];
            __builtin_memcpy(a, &v[k], sizeof(a));
#pragma unroll
            for (int j = 0; j < APU; ++j) add(a[j]);
        }
    } else {
        for (int64_t e = c0 + threadIdx.x; e < dim; e += kThreads) add(sv[e]);
    }
    const double t = block_sum(acc, red);
    const uint64_t q = block_sum_u64(qacc, qred);
    if (threadIdx.x == 0) {
        blk[blockIdx.x] = t;
        qblk[blockIdx.x] = q;
    }
}

// Per shot: target T = floor(u * P_total) (P_total = qblk[nb]; u clamped to
// [0, 1)), its chunk (the first c with qblk[c + 1] > T, qblk holding the
// exclusive prefixes and the total at nb) and the chunk's shot count.
__global__ void k_sample_locate(const double *u, int64_t shots, const uint64_t *qblk, int64_t nb, uint64_t *tgt,
                                uint32_t *chunk, uint32_t *counts) {
    const uint64_t total = qblk[nb];
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < shots; s += (int64_t)gridDim.x * blockDim.x) {
        double x = u[s];
        x = x > 0.0 ? (x < 1.0 ? x : 0.99999999999999989) : 0.0;  // NaN -> 0
        uint64_t t = (uint64_t)(x * (double)total);
        if (t >= total) t = total - 1;
        int64_t lo = 0, hi = nb - 1;  // first c with incl(c) > t; incl(nb - 1) = total > t
        while (lo < hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            if (qblk[mid + 1] > t) hi = mid;
            else lo = mid + 1;
        }
        tgt[s] = t;
        chunk[s] = (uint32_t)lo;
        atomicAdd(&counts[lo], 1u);
    }
}

// bucket[offsets[c] + k] = the k-th shot (in arrival order) of chunk c;
// cursor[] starts at zero.
__global__ void k_sample_scatter(const uint32_t *chunk, int64_t shots, const uint64_t *offsets, uint32_t *cursor,
                                 int64_t *bucket) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < shots; s += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t c = chunk[s];
        bucket[offsets[c] + atomicAdd(&cursor[c], 1u)] = s;
    }
}

// One workgroup per chunk: the chunk's inclusive integer prefix in LDS
// (lane t sums its 16 consecutive weights, a workgroup scan, then the
// per-element prefixes), then for each of the chunk's shots the first index
// whose prefix (plus the chunk's offset) exceeds the shot's target.
template <typename T>
__global__ void __launch_bounds__(kThreads) k_sample_chunks(const typename VT<T>::c *sv, int64_t dim,
                                                            const uint64_t *qblk, const uint64_t *offsets,
                                                            const int64_t *bucket, const uint64_t *tgt,
                                                            int64_t *out) {
    constexpr int per = kScanChunk / kThreads;
    __shared__ uint64_t ts[kThreads];
    __shared__ uint64_t buf[kScanChunk + kScanChunk / per];
    const uint64_t b0 = offsets[blockIdx.x], b1 = offsets[blockIdx.x + 1];
    if (b0 == b1) return;  // no shot in this chunk (uniform over the workgroup)
    auto at = [](int e) { return e + e / per; };
    const int64_t c0 = (int64_t)blockIdx.x * kScanChunk;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int e = k * kThreads + t;
        buf[at(e)] = c0 + e < dim ? cdf_weight(sv[c0 + e], kSampleScale) : 0ull;
    }
    __syncthreads();
    uint64_t v[per], run = 0;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        run += buf[at(t * per + j)];
        v[j] = run;
    }
    const uint64_t incl = block_inclusive_scan_u64(run, ts);
    const uint64_t base = incl - run + qblk[blockIdx.x];
    __syncthreads();  // every lane has read its elements
#pragma unroll
    for (int j = 0; j < per; ++j) buf[at(t * per + j)] = base + v[j];
    __syncthreads();
    const int64_t n_here = dim - c0 < kScanChunk ? dim - c0 : kScanChunk;
    for (uint64_t b = b0 + t; b < b1; b += kThreads) {
        const int64_t s = bucket[b];
        const uint64_t x = tgt[s];
        int lo = 0, hi = (int)n_here - 1;  // first i with prefix(i) > x (the chunk was chosen so that one exists)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (buf[at(mid)] > x) hi = mid;
            else lo = mid + 1;
        }
        out[s] = c0 + lo;
    }
}

// Exclusive prefix sums of x[0..n) (u32 or u64) into out[0..n), out[n] =
// the total, over 4096-entry tiles in parallel: tile sums, one workgroup's
// scan of the tile sums, then every tile's scan with its offset (x may be
// out: each workgroup reads its whole tile before writing it).
constexpr int kScanTile = 4096;
template <typename X>
__global__ void __launch_bounds__(kThreads) k_tile_sums(const X *x, int64_t n, uint64_t *sums) {
    constexpr int per = kScanTile / kThreads;
    __shared__ uint64_t red[kThreads / 64];
    const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x;
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < per; ++k)
        if (b + k * kThreads < n) acc += x[b + k * kThreads];
    const uint64_t t = block_sum_u64(acc, red);
    if (threadIdx.x == 0) sums[blockIdx.x] = t;
}
template <typename X>
__global__ void __launch_bounds__(kThreads) k_tile_scan(const X *x, int64_t n, const uint64_t *offs, int64_t ntiles,
                                                        uint64_t *out) {
    constexpr int per = kScanTile / kThreads;
    __shared__ uint64_t ts[kThreads];
    __shared__ uint64_t buf[kScanTile + kScanTile / per];
    auto at = [](int e) { return e + e / per; };
    const int64_t c0 = (int64_t)blockIdx.x * kScanTile;
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int e = k * kThreads + t;
        buf[at(e)] = c0 + e < n ? (uint64_t)x[c0 + e] : 0ull;
    }
    __syncthreads();
    uint64_t v[per], run = 0;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        v[j] = run;  // exclusive within the lane's 16
        run += buf[at(t * per + j)];
    }
    const uint64_t incl = block_inclusive_scan_u64(run, ts);
    const uint64_t base = incl - run + offs[blockIdx.x];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < per; ++j) buf[at(t * per + j)] = base + v[j];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < per; ++k) {
        const int e = k * kThreads + t;
        if (c0 + e < n) out[c0 + e] = buf[at(e)];
    }
    if (blockIdx.x == ntiles - 1 && t == kThreads - 1) out[n] = offs[ntiles];
}
// out[w] = sum of x over 4096-entry tile w (a fixed summation order)
__global__ void __launch_bounds__(kThreads) k_tile_sums_f64(const double *x, int64_t n, double *out) {
    constexpr int per = kScanTile / kThreads;
    __shared__ double red[kThreads / 64];
    const int64_t b = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x;
    double acc = 0;
#pragma unroll
    for (int k = 0; k < per; ++k)
        if (b + k * kThreads < n) acc += x[b + k * kThreads];
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}
// sum of x[0..n) into *out (one workgroup, coalesced)
__global__ void __launch_bounds__(kThreads) k_sum_f64(const double *x, int64_t n, double *out) {
    __shared__ double red[kThreads / 64];
    double acc = 0;
    for (int64_t i = threadIdx.x; i < n; i += kThreads) acc += x[i];
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) *out = t;
}
// tiles: (n + kScanTile - 1) / kScanTile <= kThreads * 1024; sums holds
// tiles + 1 entries
template <typename X>
hipError_t launch_exclusive_scan(const X *x, int64_t n, uint64_t *sums, uint64_t *out, hipStream_t s) {
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles < 1 || tiles > 0x7fffffffll) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_tile_sums<X>), dim3((unsigned)tiles), dim3(kThreads), 0, s, x, n, sums);
    hipLaunchKernelGGL(k_scan_qblocks, dim3(1), dim3(kThreads), 0, s, sums, tiles);
    hipLaunchKernelGGL((k_tile_scan<X>), dim3((unsigned)tiles), dim3(kThreads), 0, s, x, n, (const uint64_t *)sums,
                       tiles, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_sample_totals(const void *sv, int64_t dim, double *blk, uint64_t *qblk, uint64_t *sums,
                                hipStream_t s) {
    const int64_t nb = (dim + kScanChunk - 1) / kScanChunk;
    if (nb > 0x7fffffffll) return hipErrorInvalidValue;
    using C = typename VT<T>::c;
    hipLaunchKernelGGL((k_sample_totals<T>), dim3((unsigned)nb), dim3(kThreads), 0, s, (const C *)sv, dim, blk, qblk);
    // the FP64 total through per-tile sums (the scan workspace holds them
    // until the fixed-point scan below needs it)
    const int64_t tiles = (nb + kScanTile - 1) / kScanTile;
    double *tsum = reinterpret_cast<double *>(sums);
    hipLaunchKernelGGL(k_tile_sums_f64, dim3((unsigned)tiles), dim3(kThreads), 0, s, (const double *)blk, nb, tsum);
    hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(kThreads), 0, s, (const double *)tsum, tiles, blk + nb);
    return launch_exclusive_scan<uint64_t>(qblk, nb, sums, qblk, s);
}

template <typename T>
hipError_t launch_sample_chunks(const void *sv, int64_t dim, const uint64_t *qblk, const double *u, int64_t shots,
                                SampleScratch w, int64_t *out, hipStream_t s) {
    const int64_t nb = (dim + kScanChunk - 1) / kScanChunk;
    if (nb > 0x7fffffffll || shots <= 0) return shots <= 0 ? hipSuccess : hipErrorInvalidValue;
    using C = typename VT<T>::c;
    const int64_t grid = std::min<int64_t>((shots + kThreads - 1) / kThreads, 1 << 14);
    hipError_t e = hipMemsetAsync(w.counts, 0, (size_t)nb * sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sample_locate, dim3((unsigned)grid), dim3(kThreads), 0, s, u, shots, qblk, nb, w.tgt,
                       w.chunk, w.counts);
    e = launch_exclusive_scan<uint32_t>(w.counts, nb, w.sums, w.offsets, s);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(w.counts, 0, (size_t)nb * sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sample_scatter, dim3((unsigned)grid), dim3(kThreads), 0, s, (const uint32_t *)w.chunk, shots,
                       (const uint64_t *)w.offsets, w.counts, w.bucket);
    hipLaunchKernelGGL((k_sample_chunks<T>), dim3((unsigned)nb), dim3(kThreads), 0, s, (const C *)sv, dim, qblk,
                       (const uint64_t *)w.offsets, (const int64_t *)w.bucket, (const uint64_t *)w.tgt, out);
    return hipGetLastError();
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

// In place: the 2x2 matrix m on index bit `bit` (state precision, the same
// products as the streaming kernel), and, if pbit >= 0, the FP64 populations
// of index bit pbit in the updated state: partials[blk] (bit = 0) and
// partials[nblk + blk] (bit = 1).
template <typename T>
__global__ void __launch_bounds__(kThreads) k_apply1q_probs(typename VT<T>::c *sv, int64_t dim, int bit, double2 m00,
                                                            double2 m01, double2 m10, double2 m11, int pbit,
                                                            double *partials, int nblk) {
    using C = typename VT<T>::c;
    __shared__ double red[kThreads / 64];
    C a00, a01, a10, a11;
    a00.x = (T)m00.x, a00.y = (T)m00.y, a01.x = (T)m01.x, a01.y = (T)m01.y;
    a10.x = (T)m10.x, a10.y = (T)m10.y, a11.x = (T)m11.x, a11.y = (T)m11.y;
    const int64_t half = dim >> 1, tb = 1ll << bit;
    double s0 = 0, s1 = 0;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < half; u += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = ((u >> bit) << (bit + 1)) | (u & (tb - 1)), y = x | tb;
        const C v0 = sv[x], v1 = sv[y];
        const C w0 = cfma(a01, v1, cmul(a00, v0)), w1 = cfma(a11, v1, cmul(a10, v0));
        sv[x] = w0;
        sv[y] = w1;
        if (pbit >= 0) {
            const double2 d0 = to_d(w0), d1 = to_d(w1);
            const double q0 = d0.x * d0.x + d0.y * d0.y, q1 = d1.x * d1.x + d1.y * d1.y;
            if ((x >> pbit) & 1) s1 += q0; else s0 += q0;
            if ((y >> pbit) & 1) s1 += q1; else s0 += q1;
        }
    }
    if (pbit >= 0) {  // uniform
        const double t0 = block_sum(s0, red), t1 = block_sum(s1, red);
        if (threadIdx.x == 0) {
            partials[blockIdx.x] = t0;
            partials[nblk + blockIdx.x] = t1;
        }
    }
}

// Complex64 form of k_apply1q_probs with 16-byte accesses.  B0 (bit = 0): a
// unit is one float4, the pair (x, x | 1).  Otherwise a unit is the float4s
// at x and x | 2^bit (two neighbouring pairs).  Each wave walks GPT
// consecutive 64-unit runs per iteration, so every access instruction of a
// wave is one contiguous run; the products per amplitude are those of the
// generic kernel and the populations are accumulated in FP64 the same way.
template <bool B0, bool NT>
#ifndef DESV_A1Q_GPT
#define DESV_A1Q_GPT 8
#endif
__global__ void __launch_bounds__(kThreads) k_apply1q_probs_c64(float4 *sv, int64_t units, int bit, float2 a00,
                                                                float2 a01, float2 a10, float2 a11, int pbit,
                                                                double *partials, int nblk, int xcd) {
    constexpr int GPT = DESV_A1Q_GPT;
    __shared__ double red[kThreads / 64];
    const int64_t tb = 1ll << bit;
    const int lane = threadIdx.x & 63;
    // workgroups are dealt to the XCDs round-robin: XCD x takes the x-th
    // slice of the runs (one L2 per XCD sees one contiguous range)
    int64_t bid = blockIdx.x;
    if (xcd > 1) {
        const int64_t per = (int64_t)gridDim.x / xcd;
        if (bid < per * xcd) bid = (bid % xcd) * per + bid / xcd;
    }
    const int64_t wave0 = (bid * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    double s0 = 0, s1 = 0;
    auto pop = [&](int64_t e, float2 w) {
        const double2 d = to_d(w);
        const double q = d.x * d.x + d.y * d.y;
        if ((e >> pbit) & 1) s1 += q; else s0 += q;
    };
    for (int64_t w = wave0; w * (GPT * 64) < units; w += nwaves) {
        int64_t ix[GPT], iy[GPT];
        float4 v[GPT], u[GPT];
        bool on[GPT];
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
            const int64_t un = (w * GPT + j) * 64 + lane;
            on[j] = un < units;
            if constexpr (B0) {
                ix[j] = un;
                iy[j] = un;
                v[j] = ld<float4, NT>(sv + (on[j] ? un : 0));  // unconditional: no per-lane branch
            } else {
                const int64_t a = 2 * un, x = ((a >> bit) << (bit + 1)) | (a & (tb - 1));
                ix[j] = x >> 1;
                iy[j] = (x | tb) >> 1;
                v[j] = ld<float4, NT>(sv + (on[j] ? ix[j] : 0));
                u[j] = ld<float4, NT>(sv + (on[j] ? iy[j] : 0));
            }
        }
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
            if (!on[j]) continue;
            if constexpr (B0) {
                const float2 v0 = make_float2(v[j].x, v[j].y), v1 = make_float2(v[j].z, v[j].w);
                const float2 w0 = cfma(a01, v1, cmul(a00, v0)), w1 = cfma(a11, v1, cmul(a10, v0));
                st<float4, NT>(sv + ix[j], make_float4(w0.x, w0.y, w1.x, w1.y));
                if (pbit >= 0) {
                    pop(2 * ix[j], w0);
                    pop(2 * ix[j] + 1, w1);
                }
            } else {
                const float2 p0 = make_float2(v[j].x, v[j].y), p1 = make_float2(v[j].z, v[j].w);
                const float2 q0 = make_float2(u[j].x, u[j].y), q1 = make_float2(u[j].z, u[j].w);
                const float2 w0 = cfma(a01, q0, cmul(a00, p0)), y0 = cfma(a11, q0, cmul(a10, p0));
                const float2 w1 = cfma(a01, q1, cmul(a00, p1)), y1 = cfma(a11, q1, cmul(a10, p1));
                st<float4, NT>(sv + ix[j], make_float4(w0.x, w0.y, w1.x, w1.y));
                st<float4, NT>(sv + iy[j], make_float4(y0.x, y0.y, y1.x, y1.y));
                if (pbit >= 0) {
                    pop(2 * ix[j], w0);
                    pop(2 * ix[j] + 1, w1);
                    pop(2 * iy[j], y0);
                    pop(2 * iy[j] + 1, y1);
                }
            }
        }
    }
    if (pbit >= 0) {  // uniform
        const double t0 = block_sum(s0, red), t1 = block_sum(s1, red);
        if (threadIdx.x == 0) {
            partials[blockIdx.x] = t0;
            partials[nblk + blockIdx.x] = t1;
        }
    }
}

#ifndef DESV_A1Q_XCD
#define DESV_A1Q_XCD direct_xcd()
#endif
template <typename T>
hipError_t launch_apply1q_probs(void *sv, int64_t dim, int bit, const double *m, int pbit, double *partials, int nblk,
                                hipStream_t s) {
    if constexpr (sizeof(T) == 4) {
        // 16-byte units from 2^10 amplitudes up (every unit count a multiple of 256)
        if (dim >= 1024) {
            const float2 a00 = make_float2((float)m[0], (float)m[1]), a01 = make_float2((float)m[2], (float)m[3]);
            const float2 a10 = make_float2((float)m[4], (float)m[5]), a11 = make_float2((float)m[6], (float)m[7]);
            float4 *p = (float4 *)sv;
            if (bit == 0)
                hipLaunchKernelGGL((k_apply1q_probs_c64<true, false>), dim3(nblk), dim3(kThreads), 0, s, p, dim / 2,
                                   bit, a00, a01, a10, a11, pbit, partials, nblk, DESV_A1Q_XCD);
            else if (direct_shape(false, true, bit).nt)
                hipLaunchKernelGGL((k_apply1q_probs_c64<false, true>), dim3(nblk), dim3(kThreads), 0, s, p, dim / 4,
                                   bit, a00, a01, a10, a11, pbit, partials, nblk, DESV_A1Q_XCD);
            else
                hipLaunchKernelGGL((k_apply1q_probs_c64<false, false>), dim3(nblk), dim3(kThreads), 0, s, p, dim / 4,
                                   bit, a00, a01, a10, a11, pbit, partials, nblk, DESV_A1Q_XCD);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_apply1q_probs<T>), dim3(nblk), dim3(kThreads), 0, s, (typename VT<T>::c *)sv, dim, bit,
                       make_double2(m[0], m[1]), make_double2(m[2], m[3]), make_double2(m[4], m[5]),
                       make_double2(m[6], m[7]), pbit, partials, nblk);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Layers of single-qubit monomial operators (noise trajectories, reference
// noise/kraus_channels.py:104-240): out[x] = coef(x) in[x ^ F], one streaming
// pass in place, and the FP64 populations of the result over up to 12 bits.
// A lane owns vector v and, when F reaches past the vector, its partner
// v ^ fv: both are read before either is written.  Index bits below t.lamp
// (amplitude units) are fixed per lane for the whole launch; the bits above
// are uniform across a wave's run of 64 vectors, so their coefficients come
// from byte-indexed tables and the population bits among them from blockIdx.y.
// ---------------------------------------------------------------------------
__device__ inline uint64_t pdep64(uint64_t x, uint64_t mask) {
    uint64_t r = 0;
    for (uint64_t m = mask; m && x; m &= m - 1, x >>= 1)
        if (x & 1) r |= m & (~m + 1);
    return r;
}
__device__ inline uint32_t pext64(uint64_t x, uint64_t mask) {
    uint32_t r = 0;
    int k = 0;
    for (uint64_t m = mask; m; m &= m - 1, ++k)
        if (x & m & (~m + 1)) r |= 1u << k;
    return r;
}

template <typename T, bool PAIR, bool POPS>
__global__ void __launch_bounds__(kThreads) k_mono_layer(typename VT<T>::v *sv, const MonoTask t) {
    using V = typename VT<T>::v;
    constexpr int PV = VT<T>::per_vec, PVB = PV == 2 ? 1 : 0, NS = PAIR ? 2 : 1, GPT = 4;
    __shared__ double acc_l[POPS ? kThreads * NS * PV : 1];
    __shared__ int bin_l[POPS ? kThreads * NS * PV : 1];
    const int lane = threadIdx.x & 63;
    const int64_t vlo = (PAIR && t.hb < 6) ? (int64_t)insert_zero(lane, t.hb) : (int64_t)lane;
    const int64_t wlo = vlo ^ (t.fv & ((1ll << t.lb) - 1));
    // per-lane factors of the low bits, and the population bins of the lane part
    double2 cl[NS * PV];
    int bl[NS * PV];
    double acc[NS * PV];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int e = 0; e < PV; ++e) {
            const int64_t a = ((s ? wlo : vlo) << PVB) | e;
            double2 c = make_double2(1.0, 0.0);
            for (int i = 0; i < t.nlow; ++i) c = cmul(c, t.lc[i][(a >> t.lbit[i]) & 1]);
            cl[s * PV + e] = c;
            acc[s * PV + e] = 0.0;
            bl[s * PV + e] = POPS ? (int)pext64((uint64_t)a | (s ? t.hbit_amp : 0), t.plmask) : 0;
        }
    const uint64_t hy = pdep64(blockIdx.y, t.phmask);
    const int64_t wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
    const int64_t nw = (int64_t)gridDim.x * (kThreads / 64);
    const uint64_t nr = ~t.rmask;
    const uint64_t step = pdep64((uint64_t)(nw * GPT), t.rmask);
    uint64_t hd = pdep64((uint64_t)(wave * GPT), t.rmask);
    for (int64_t r0 = wave * GPT; r0 < t.nruns; r0 += nw * GPT, hd = ((hd | nr) + step) & t.rmask) {
        int64_t iv[GPT];
        uint64_t hu[GPT];  // the run's bits >= lb: uniform across the wave
        bool on[GPT];
        V a[GPT], b[GPT];
        uint64_t h = hd;
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
            hu[j] = h | hy;
            iv[j] = (int64_t)hu[j] | vlo;
            on[j] = r0 + j < t.nruns && iv[j] < t.nvec;
            // unconditional loads (an index off the state reads vector 0 and
            // its partner), so that no branch separates them
            const int64_t ia = on[j] ? iv[j] : 0;
            a[j] = sv[ia];
            if constexpr (PAIR) b[j] = sv[ia ^ t.fv];
            h = ((h | nr) + t.rstep1) & t.rmask;
        }
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
            double2 ch[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                // bits >= lamp of the amplitude index = bits >= lb of the vector index
                const uint64_t hv = (s ? (hu[j] ^ (uint64_t)t.fv) : hu[j]) >> t.lb;
                double2 c = make_double2(1.0, 0.0);
                for (int k = 0; k < t.ntab; ++k)
                    c = cmul(c, t.tab[k * 256 + __builtin_amdgcn_readfirstlane((int)((hv >> (8 * k)) & 255))]);
                ch[s] = c;
            }
            V outv[NS];
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                // the v outputs read the partner (or v itself), the w outputs read v
                T x[2 * PV], o[2 * PV];
                __builtin_memcpy(x, (PAIR && s == 0) ? &b[j] : &a[j], sizeof(V));
#pragma unroll
                for (int e = 0; e < PV; ++e) {
                    const double2 c = cmul(ch[s], cl[s * PV + e]);
                    const bool sw = PV == 2 && t.f0;
                    const T xr = sw ? x[2 * (e ^ (PV - 1))] : x[2 * e], xi = sw ? x[2 * (e ^ (PV - 1)) + 1] : x[2 * e + 1];
                    const T cr = (T)c.x, ci = (T)c.y;
                    o[2 * e] = cr * xr - ci * xi;
                    o[2 * e + 1] = cr * xi + ci * xr;
                    if constexpr (POPS) {
                        const double dr = (double)o[2 * e], di = (double)o[2 * e + 1];
                        acc[s * PV + e] += on[j] ? dr * dr + di * di : 0.0;
                    }
                }
                __builtin_memcpy(&outv[s], o, sizeof(V));
            }
            if (t.store && on[j]) {
                sv[iv[j]] = outv[0];
                if constexpr (PAIR) sv[iv[j] ^ t.fv] = outv[NS - 1];
            }
        }
    }
    if constexpr (POPS) {
#pragma unroll
        for (int i = 0; i < NS * PV; ++i) {
            acc_l[i * kThreads + threadIdx.x] = acc[i];
            bin_l[i * kThreads + threadIdx.x] = bl[i];
        }
        __syncthreads();
        // partials[((y << npl | b) * gridDim.x + blockIdx.x) * NS + s]: the w
        // half of this workgroup lands in the bins of y ^ fph
        const int nb = 1 << t.npl;
        for (int q = threadIdx.x; q < nb * NS; q += kThreads) {
            const int s = q / nb, bb = q % nb;
            double sum = 0.0;
            for (int e = 0; e < PV; ++e)
                for (int i = 0; i < kThreads; ++i) {
                    const int slot = (s * PV + e) * kThreads + i;
                    if (bin_l[slot] == bb) sum += acc_l[slot];
                }
            const int64_t y = s ? (blockIdx.y ^ t.fph) : blockIdx.y;
            t.partials[((((int64_t)y << t.npl) | bb) * gridDim.x + blockIdx.x) * NS + s] = sum;
        }
    }
}

template <typename T>
hipError_t launch_mono_layer(void *sv, const MonoTask &t, int gx, int gy, hipStream_t s) {
    using V = typename VT<T>::v;
    const dim3 grid(gx, gy), block(kThreads);
    V *p = (V *)sv;
    const bool pops = t.partials != nullptr;
    if (t.fv != 0) {
        if (pops) hipLaunchKernelGGL((k_mono_layer<T, true, true>), grid, block, 0, s, p, t);
        else hipLaunchKernelGGL((k_mono_layer<T, true, false>), grid, block, 0, s, p, t);
    } else {
        if (pops) hipLaunchKernelGGL((k_mono_layer<T, false, true>), grid, block, 0, s, p, t);
        else hipLaunchKernelGGL((k_mono_layer<T, false, false>), grid, block, 0, s, p, t);
    }
    return hipGetLastError();
}

// Complex64 forms of k_norm2 (PROBS = false) and k_qubit_probs (PROBS =
// true) with 16-byte loads: a float4 holds amplitudes 2f and 2f + 1, and each
// wave reads 4 contiguous 64-float4 runs per iteration, so every load
// instruction of a wave is one contiguous run and four are in flight per
// lane.  The FP64 terms per amplitude are those of the generic kernels.
template <bool PROBS>
__global__ void __launch_bounds__(kThreads) k_pops_c64(const float4 *sv, int64_t nf4, int bit, double *partials) {
    constexpr int GPT = 4;
    __shared__ double red[kThreads / 64];
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    double p0 = 0, p1 = 0;
    for (int64_t w = wave0; w * (GPT * 64) < nf4; w += nwaves) {
        float4 v[GPT];
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
            const int64_t f = (w * GPT + j) * 64 + lane;
            v[j] = f < nf4 ? sv[f] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
            const int64_t e = 2 * ((w * GPT + j) * 64 + lane);
            const double2 a = to_d(make_float2(v[j].x, v[j].y)), b = to_d(make_float2(v[j].z, v[j].w));
            const double pa = a.x * a.x + a.y * a.y, pb = b.x * b.x + b.y * b.y;
            if constexpr (PROBS) {
                if ((e >> bit) & 1) p1 += pa; else p0 += pa;
                if (((e + 1) >> bit) & 1) p1 += pb; else p0 += pb;
            } else {
                p0 += pa;
                p0 += pb;
            }
        }
    }
    const double s0 = block_sum(p0, red);
    if constexpr (PROBS) {
        const double s1 = block_sum(p1, red);
        if (threadIdx.x == 0) {
            partials[blockIdx.x] = s0;
            partials[gridDim.x + blockIdx.x] = s1;
        }
    } else if (threadIdx.x == 0) {
        partials[blockIdx.x] = s0;
    }
}

template <typename T>
hipError_t launch_norm2(const void *sv, int64_t dim, double *partials, int nblk, hipStream_t s) {
    if (sizeof(T) == 4 && dim >= 2)
        hipLaunchKernelGGL((k_pops_c64<false>), dim3(nblk), dim3(kThreads), 0, s, (const float4 *)sv, dim / 2, 0,
                           partials);
    else
        hipLaunchKernelGGL((k_norm2<T>), dim3(nblk), dim3(kThreads), 0, s, (const typename VT<T>::c *)sv, dim,
                           partials);
    return hipGetLastError();
}
// Complex64 form of k_inner: 16-byte loads of both vectors (a float4 is
// amplitudes 2f and 2f + 1), 4 contiguous 64-float4 runs per wave and
// iteration; the FP64 terms per amplitude are those of the generic kernel.
__global__ void __launch_bounds__(kThreads) k_inner_c64(const float4 *a, const float4 *b, int64_t nf4,
                                                        double *partials) {
    constexpr int GPT = 4;
    __shared__ double red[kThreads / 64];
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    double re = 0, im = 0;
    for (int64_t w = wave0; w * (GPT * 64) < nf4; w += nwaves) {
        float4 x[GPT], y[GPT];
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
            const int64_t f = (w * GPT + j) * 64 + lane;
            const bool on = f < nf4;
            x[j] = on ? a[f] : make_float4(0.f, 0.f, 0.f, 0.f);
            y[j] = on ? b[f] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < GPT; ++j) {
            const double2 x0 = to_d(make_float2(x[j].x, x[j].y)), x1 = to_d(make_float2(x[j].z, x[j].w));
            const double2 y0 = to_d(make_float2(y[j].x, y[j].y)), y1 = to_d(make_float2(y[j].z, y[j].w));
            re += x0.x * y0.x + x0.y * y0.y;
            im += x0.x * y0.y - x0.y * y0.x;
            re += x1.x * y1.x + x1.y * y1.y;
            im += x1.x * y1.y - x1.y * y1.x;
        }
    }
    const double sr = block_sum(re, red);
    const double si = block_sum(im, red);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = sr;
        partials[gridDim.x + blockIdx.x] = si;
    }
}

template <typename T>
hipError_t launch_inner(const void *a, const void *b, int64_t dim, double *partials, int nblk, hipStream_t s) {
    if (sizeof(T) == 4 && dim >= 2)
        hipLaunchKernelGGL(k_inner_c64, dim3(nblk), dim3(kThreads), 0, s, (const float4 *)a, (const float4 *)b,
                           dim / 2, partials);
    else
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
    if (sizeof(T) == 4 && dim >= 2)
        hipLaunchKernelGGL((k_pops_c64<true>), dim3(nblk), dim3(kThreads), 0, s, (const float4 *)sv, dim / 2, bit,
                           partials);
    else
        hipLaunchKernelGGL((k_qubit_probs<T>), dim3(nblk), dim3(kThreads), 0, s, (const typename VT<T>::c *)sv, dim,
                           bit, partials);
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
hipError_t launch_dense_matvec(void *out, const void *in, int n, const double2 *H, hipStream_t s) {
    hipLaunchKernelGGL((k_dense_matvec<T>), dim3(ew_grid(1ll << n)), dim3(kThreads), 0, s,
                       (typename VT<T>::c *)out, (const typename VT<T>::c *)in, n, H);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Krylov bases (block Gram-Schmidt and Ritz rotation for Lanczos): streaming,
// HBM-bound; FP64 accumulation whatever the state precision.
// ---------------------------------------------------------------------------
// <v_j|w> for j in the workgroup row's group of kInnerGroup vectors: w is read
// once per group, each v_j once.  Partials: [(2 j + {re, im}) * nblk + block].
template <typename T>
__global__ void __launch_bounds__(kThreads) k_inner_many(const typename VT<T>::c *w, BasisPtrs vs, int m,
                                                         int64_t dim, double *partials) {
    using C = typename VT<T>::c;
    __shared__ double red[kThreads / 64];
    const int j0 = blockIdx.y * kInnerGroup;
    const int cnt = min(kInnerGroup, m - j0);
    double re[kInnerGroup], im[kInnerGroup];
#pragma unroll
    for (int g = 0; g < kInnerGroup; ++g) re[g] = im[g] = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 y = to_d(w[i]);
#pragma unroll
        for (int g = 0; g < kInnerGroup; ++g) {
            if (g < cnt) {
                const double2 x = to_d(static_cast<const C *>(vs.p[j0 + g])[i]);
                re[g] += x.x * y.x + x.y * y.y;
                im[g] += x.x * y.y - x.y * y.x;
            }
        }
    }
    for (int g = 0; g < cnt; ++g) {
        const double sr = block_sum(re[g], red);
        const double si = block_sum(im[g], red);
        if (threadIdx.x == 0) {
            partials[(int64_t)(2 * (j0 + g)) * gridDim.x + blockIdx.x] = sr;
            partials[(int64_t)(2 * (j0 + g) + 1) * gridDim.x + blockIdx.x] = si;
        }
    }
}

// out_l[i] = sum_j coef[j * p + l] * in_j[i] for l < p.  Every input element
// of index i is loaded before any output element of index i is stored, by the
// same lane, so outputs may alias inputs (in-place basis rotation).
template <typename T, int M>
__global__ void __launch_bounds__(kThreads) k_combine(BasisPtrs in, BasisPtrs out, int m, int p,
                                                      const double2 *__restrict__ coef, int64_t dim) {
    using C = typename VT<T>::c;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        double2 x[M];
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (j < m) x[j] = to_d(static_cast<const C *>(in.p[j])[i]);
        for (int l = 0; l < p; ++l) {
            double2 acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int j = 0; j < M; ++j)
                if (j < m) acc = cfma(coef[j * p + l], x[j], acc);
            C r;
            r.x = (T)acc.x;
            r.y = (T)acc.y;
            static_cast<C *>(const_cast<void *>(out.p[l]))[i] = r;
        }
    }
}

// The same for at most 2 outputs, streaming over the inputs: each workgroup
// takes E x 256 consecutive elements (E per lane, 256 apart) of every input
// in turn, so each input is read in E KiB (complex64) or 2E KiB (complex128)
// contiguous runs per workgroup instead of 256 elements, and the lane keeps
// only its outputs' accumulators.  An output may alias an input: the lane
// loads all of its elements' inputs before storing their outputs.
template <typename T, int E>
__global__ void __launch_bounds__(kThreads) k_combine_stream(BasisPtrs in, BasisPtrs out, int m, int p,
                                                             const double2 *__restrict__ coef, int64_t dim) {
    using C = typename VT<T>::c;
    const int64_t tile = (int64_t)E * kThreads;
    for (int64_t t0 = (int64_t)blockIdx.x * tile; t0 < dim; t0 += (int64_t)gridDim.x * tile) {
        double2 acc[2][E];
#pragma unroll
        for (int e = 0; e < E; ++e) acc[0][e] = acc[1][e] = make_double2(0.0, 0.0);
        for (int j = 0; j < m; ++j) {
            const C *src = static_cast<const C *>(in.p[j]);
            const double2 c0 = coef[j * p], c1 = p > 1 ? coef[j * p + 1] : make_double2(0.0, 0.0);
            double2 x[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int64_t i = t0 + e * kThreads + threadIdx.x;
                x[e] = i < dim ? to_d(src[i]) : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                acc[0][e] = cfma(c0, x[e], acc[0][e]);
                acc[1][e] = cfma(c1, x[e], acc[1][e]);
            }
        }
        for (int l = 0; l < p; ++l)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int64_t i = t0 + e * kThreads + threadIdx.x;
                if (i < dim) {
                    C r;
                    r.x = (T)acc[l][e].x;
                    r.y = (T)acc[l][e].y;
                    static_cast<C *>(const_cast<void *>(out.p[l]))[i] = r;
                }
            }
    }
}

// Amplitude i = (u0 - 1/2) + i (u1 - 1/2), u from SplitMix64 of (seed, i):
// a reproducible start vector with no host transfer (not normalised).
__device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void k_fill_random(typename VT<T>::c *sv, int64_t dim, uint64_t seed) {
    using C = typename VT<T>::c;
    const uint64_t key = splitmix64(seed);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < dim; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t a = splitmix64(key ^ (2 * (uint64_t)i));
        const uint64_t b = splitmix64(key ^ (2 * (uint64_t)i + 1));
        C r;
        r.x = (T)((double)(a >> 11) * 0x1.0p-53 - 0.5);
        r.y = (T)((double)(b >> 11) * 0x1.0p-53 - 0.5);
        sv[i] = r;
    }
}

template <typename T>
hipError_t launch_inner_many(const void *w, const BasisPtrs &vs, int m, int64_t dim, double *partials, int nblk,
                             hipStream_t s) {
    const int rows = (m + kInnerGroup - 1) / kInnerGroup;
    hipLaunchKernelGGL((k_inner_many<T>), dim3(nblk, rows), dim3(kThreads), 0, s, (const typename VT<T>::c *)w, vs,
                       m, dim, partials);
    return hipGetLastError();
}

#ifndef DESV_COMBINE_E
#define DESV_COMBINE_E 2
#endif
template <typename T>
hipError_t launch_combine(const BasisPtrs &in, const BasisPtrs &out, int m, int p, const double2 *coef, int64_t dim,
                          hipStream_t s) {
    const dim3 g(ew_grid(dim)), b(kThreads);
    if (p <= 2 && DESV_COMBINE_E > 0) {
        constexpr int E = DESV_COMBINE_E > 0 ? DESV_COMBINE_E : 1;
        const int64_t tiles = (dim + (int64_t)E * kThreads - 1) / ((int64_t)E * kThreads);
        const dim3 gs((unsigned)std::min<int64_t>(std::max<int64_t>(tiles, 1), 65536));
        hipLaunchKernelGGL((k_combine_stream<T, E>), gs, b, 0, s, in, out, m, p, coef, dim);
    } else if (m <= 8)
        hipLaunchKernelGGL((k_combine<T, 8>), g, b, 0, s, in, out, m, p, coef, dim);
    else if (m <= 16)
        hipLaunchKernelGGL((k_combine<T, 16>), g, b, 0, s, in, out, m, p, coef, dim);
    else if (m <= 32)
        hipLaunchKernelGGL((k_combine<T, 32>), g, b, 0, s, in, out, m, p, coef, dim);
    else
        hipLaunchKernelGGL((k_combine<T, kMaxBasis>), g, b, 0, s, in, out, m, p, coef, dim);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_random(void *sv, int64_t dim, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL((k_fill_random<T>), dim3(ew_grid(dim)), dim3(kThreads), 0, s, (typename VT<T>::c *)sv, dim,
                       seed);
    return hipGetLastError();
}

#define DESV_INST(T)                                                                                            \
    template hipError_t launch_direct<T>(const void *, void *, int, const DirectOp &, const void *, hipStream_t);            \
    template hipError_t launch_stage<T>(int, int, bool, const StageTask *, const int32_t *, const StageTask &,  \
                                        int64_t, int, hipStream_t);                                            \
    template hipError_t launch_matvec_tasks<T>(const MatvecTask *, int, int64_t, hipStream_t);                 \
    template hipError_t launch_fill_tasks<T>(const FillTask *, int, int64_t, hipStream_t);                     \
    template hipError_t launch_fill_basis<T>(void *, int64_t, int64_t, hipStream_t);                          \
    template hipError_t launch_scale<T>(void *, int64_t, double, double, hipStream_t);                        \
    template hipError_t launch_norm2<T>(const void *, int64_t, double *, int, hipStream_t);                   \
    template hipError_t launch_inner<T>(const void *, const void *, int64_t, double *, int, hipStream_t);     \
    template hipError_t launch_probs<T>(const void *, int64_t, double *, hipStream_t);                        \
    template hipError_t launch_cdf<T>(const void *, int64_t, double *, double *, hipStream_t);                \
    template hipError_t launch_sample_totals<T>(const void *, int64_t, double *, uint64_t *, uint64_t *,       \
                                                hipStream_t);                                                  \
    template hipError_t launch_sample_chunks<T>(const void *, int64_t, const uint64_t *, const double *,      \
                                                int64_t, SampleScratch, int64_t *, hipStream_t);               \
    template hipError_t launch_qubit_probs<T>(const void *, int64_t, int, double *, int, hipStream_t);        \
    template hipError_t launch_mono_layer<T>(void *, const MonoTask &, int, int, hipStream_t);              \
    template hipError_t launch_collapse<T>(void *, int64_t, int, int, double, hipStream_t);                   \
    template hipError_t launch_pauli_expect<T>(const void *, int, const PauliItem *, int, const uint64_t *,   \
                                               double *, int, hipStream_t);                                   \
    template hipError_t launch_dense_matvec<T>(void *, const void *, int, const double2 *, hipStream_t);        \
    template hipError_t launch_inner_many<T>(const void *, const BasisPtrs &, int, int64_t, double *, int,     \
                                             hipStream_t);                                                    \
    template hipError_t launch_combine<T>(const BasisPtrs &, const BasisPtrs &, int, int, const double2 *,     \
                                          int64_t, hipStream_t);                                              \
    template hipError_t launch_fill_random<T>(void *, int64_t, uint64_t, hipStream_t);                          \
    template hipError_t launch_apply1q_probs<T>(void *, int64_t, int, const double *, int, double *, int,        \
                                                hipStream_t);

DESV_INST(float)
DESV_INST(double)

}  // namespace desv
