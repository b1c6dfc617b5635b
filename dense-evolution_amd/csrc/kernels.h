// Device kernels of the dense statevector engine (gfx950, wave64).
//
// Every kernel works on interleaved complex amplitudes (float2 / double2).
// Reductions accumulate in FP64 and are written as per-block partials that a
// second, single-block pass sums in a fixed order, so results are
// reproducible run to run.
#pragma once

#include <cstdlib>
#include <hip/hip_runtime.h>

#include "desv_internal.h"

namespace desv {

template <typename T> struct VT;
template <> struct VT<float> {
    using c = float2;
    using v = float4;  // one 16-byte global access = 2 amplitudes
    static constexpr int per_vec = 2;
};
template <> struct VT<double> {
    using c = double2;
    using v = double2;  // one 16-byte global access = 1 amplitude
    static constexpr int per_vec = 1;
};

constexpr int kThreads = 256;
constexpr int kMaxGradLDS = 64;

// One stage of one state: which tiles, which ops, which buffers.  A launch
// runs one task (passed by value) or an array of tasks (one stage of many
// batch instances); workgroup b serves tile (b - wg_begin) of the task with
// the largest wg_begin <= b.
struct StageTask {
    DevStage st;
    int32_t n;
    int32_t lo_off;      // compensated launch: offset (complex elements) of the lo matrix pool
    void *sv0;           // psi
    void *sv1;           // lambda (adjoint reverse sweep), else unused
    const DevOp *ops;
    const void *mats;    // complex pool in the state precision
    const void *ltab;    // per-lane tables of 3-qubit MFMA ops (DevOp::lt), or null
    const int32_t *rprog;  // register-group words of this stage (RegProgram), or null
    const void *rfrag;     // this stage's first register-group fragment block
    double *partials;    // gradient partials [slot][pstride]
    int64_t pstride;     // tiles per gradient slot row
    int64_t wg_begin;
};

// Stage kernel LDS behind the tile(s): the task copy, the chunk tables
// (256 global offsets, 256 slots) and the per-wave gradient partials.
constexpr int kStageLdsTask = (int)((sizeof(StageTask) + 15) / 16 * 16);
constexpr int kStageLdsExtra = kStageLdsTask + 256 * 8 + 256 * 4 + kMaxGradLDS * (kThreads / 64) * 8;


// lambda = H psi, psi2 = psi and FP64 partials of Re<psi|lambda>, for a chunk of
// one instance's amplitudes per workgroup.
// Amplitudes per workgroup of the tiled Pauli matvec (one aligned tile in
// LDS, at most 8 per lane): 2^11 at most.
constexpr int kMatvecTileBitsC64 = 11, kMatvecTileBitsC128 = 11;
// log2 amplitudes per workgroup for an n-qubit state (DESV_MATVEC_TILE_BITS
// caps it below 11).
inline int matvec_tile_bits(int n) {
    static const int cap = [] {
        const char *e = std::getenv("DESV_MATVEC_TILE_BITS");
        const int v = e ? std::atoi(e) : kMatvecTileBitsC64;
        return v >= 1 && v <= kMatvecTileBitsC64 ? v : kMatvecTileBitsC64;
    }();
    return n < cap ? n : cap;
}

struct MatvecTask {
    int32_t n;
    int32_t ngroups;
    const void *psi;
    void *lam;
    void *psi2;          // may be null
    const void *groups;  // MatvecGroup[ngroups]
    const uint64_t *z;
    const double2 *cc;
    double *epart;       // one double per workgroup of this task
    int64_t wg_begin;
    int64_t chunk;       // amplitudes per workgroup: 2^min(n, kMatvecTileBits*)
};

// dst[0..dim) = src, or |0...0> when src is null.
struct FillTask {
    void *dst;
    const void *src;
    int64_t dim;
    int64_t wg_begin;
    int64_t chunk;
};

// *dst = sum of src[0..count); then, if stamp is set, *stamp = max(*stamp,
// device wall clock) (completion time of the result).
struct SumTask {
    const double *src;
    int64_t count;
    double *dst;
    unsigned long long *stamp;
};

// Pauli-expectation work item: one flip mask, up to kTermsPerItem terms.
constexpr int kTermsPerItem = 8;
struct PauliItem {
    uint64_t x;
    int32_t begin;
    int32_t count;
};

// Pauli matvec group: terms [begin, end) share flip mask x.
struct MatvecGroup {
    uint64_t x;
    int32_t begin;
    int32_t end;
};

// One dense (k = 1, 2) or diagonal op over a whole state, global bit
// positions; applied by the streaming direct kernel.
struct DirectOp {
    int32_t k;
    int32_t diag;
    int32_t pos[2];
    uint64_t cmask;
};

// ---- launchers (engine_kernels.hip) --------------------------------------
template <typename T>
// in == out: in place; otherwise out-of-place into a second buffer of the same size.
hipError_t launch_direct(const void *in, void *out, int n, const DirectOp &op, const void *M, hipStream_t s);
// Stage launch: `tasks` (device array) with wg_task[b] = task of workgroup b,
// or, if tasks is null, `single`.
template <typename T>
// `variant`: DevStage::heavy (0 light, 1 VALU blocks k = 3..5, 2 MFMA blocks
// k = 3, 4, 3 MFMA blocks up to k = 5), or 4: register groups (complex64,
// nv = 1, tasks carry `rprog`).  `comp`: compensated matrices (the lo pool at
// mats + lo_off; complex64 only).
hipError_t launch_stage(int nv, int variant, bool comp, const StageTask *tasks, const int32_t *wg_task,
                        const StageTask &single, int64_t grid, int max_m, hipStream_t s);
// Wave-tile stage (complex64, nv = 1): one 2^m-amplitude tile per wave held
// in registers (4 waves a workgroup), m = 11 or 12; `t.rprog` = the stage's
// wave program words, `t.rfrag` its first fragment block, `nq` its MFMA ops
// (their fragment blocks are staged in LDS).
hipError_t launch_wave_stage(int m, bool comp, int nq, const StageTask &t, hipStream_t s);
template <typename T>
hipError_t launch_matvec_tasks(const MatvecTask *tasks, int ntasks, int64_t grid, hipStream_t s);
template <typename T>
hipError_t launch_fill_tasks(const FillTask *tasks, int ntasks, int64_t grid, hipStream_t s);
hipError_t launch_sum_tasks(const SumTask *tasks, int ntasks, hipStream_t s);
// stamps[1 .. count] = 0, stamps[0] = device wall clock (start of a timed run).
hipError_t launch_mark(unsigned long long *stamps, int64_t count, hipStream_t s);
// dst <- src on the device (bytes a multiple of 16).
hipError_t launch_copy16(void *dst, const void *src, size_t bytes, hipStream_t s);
// dst <- src converted between complex64 (float) and complex128 (double).
template <typename TD, typename TS>
hipError_t launch_convert(void *dst, const void *src, int64_t dim, hipStream_t s);
template <typename T>
hipError_t launch_fill_basis(void *sv, int64_t dim, int64_t index, hipStream_t s);
template <typename T>
hipError_t launch_scale(void *sv, int64_t dim, double re, double im, hipStream_t s);
template <typename T>
hipError_t launch_norm2(const void *sv, int64_t dim, double *partials, int nblk, hipStream_t s);
template <typename T>
hipError_t launch_inner(const void *a, const void *b, int64_t dim, double *partials, int nblk, hipStream_t s);
template <typename T>
hipError_t launch_probs(const void *sv, int64_t dim, double *out, hipStream_t s);
// Normalised cumulative probabilities of a state into cdf[0..dim) (blk holds
// 2 (cdf_blocks(dim) + 1) doubles of workspace: FP64 chunk totals and their
// scan, blk[cdf_blocks(dim)] = the total before normalisation, then the
// fixed-point chunk totals), and the shot search over it.
template <typename T>
hipError_t launch_cdf(const void *sv, int64_t dim, double *cdf, double *blk, hipStream_t s);
int64_t cdf_blocks(int64_t dim);
// In place 2x2 matrix (row-major re/im, 8 doubles) on index bit `bit`; if
// pbit >= 0 also the populations of index bit pbit afterwards (2 x nblk partials).
template <typename T>
hipError_t launch_apply1q_probs(void *sv, int64_t dim, int bit, const double *m, int pbit, double *partials, int nblk,
                                hipStream_t s);
// One layer of single-qubit operators with at most one nonzero entry per row
// (noise trajectories): out[x] = coef(x) in[x ^ F], in place, and optionally
// the FP64 populations of the result.  Vector = 16 bytes (VT<T>::v); the lane
// part of an index is its bits below lb (vector units) = lamp (amplitude
// units); the rest is uniform per run of 64 vectors.
constexpr int kMonoLowMax = 8;   // factors on amplitude bits < lamp
constexpr int kMonoTables = 5;   // coefficient tables over 8 amplitude bits each, from lamp up
constexpr int kMonoMaxPops = 12;
struct MonoTask {
    int64_t nvec;        // vectors in the state
    int64_t fv;          // flip mask F above the vector's own amplitudes (vector units)
    int32_t f0;          // complex64: F's bit 0 (swaps the two amplitudes of a vector)
    int32_t hb;          // highest bit of fv, -1 if fv = 0 (a lane owns v and v ^ fv, bit hb of v clear)
    int32_t lb;          // vector bits fixed by the lane: 6, or 7 when hb < 6
    int32_t lamp;        // lb in amplitude bits
    int32_t store;       // 0: populations only (the identity layer)
    int32_t nlow;        // factors on amplitude bits < lamp
    int32_t lbit[kMonoLowMax];
    double2 lc[kMonoLowMax][2];  // coefficient of row 0, row 1
    int32_t ntab;        // tables tab[k * 256 + byte] for amplitude bits [lamp + 8 k, lamp + 8 k + 8)
    int32_t npl;         // population bits of the lane part (plmask)
    const double2 *tab;
    uint64_t rmask;      // vector bits walked by the runs (>= lb, not hb, not population bits)
    uint64_t rstep1;     // lowest bit of rmask
    int64_t nruns;       // 2^popcount(rmask)
    uint64_t phmask;     // population bits among the vector bits >= lb other than hb: blockIdx.y
    uint64_t plmask;     // population bits of the lane part (amplitude units, with hbit_amp)
    uint64_t hbit_amp;   // bit hb in amplitude units when hb >= lb (set in the partner's bins), else 0
    uint32_t fph;        // pext(fv, phmask): the partner's blockIdx.y bins are y ^ fph
    double *partials;    // null: no populations; else [(y << npl | b) * gx + x] * (1 + (fv != 0)) + s
};
template <typename T>
hipError_t launch_mono_layer(void *sv, const MonoTask &t, int gx, int gy, hipStream_t s);
hipError_t launch_sample_search(const double *cdf, int64_t dim, const double *u, int64_t shots, int64_t *out,
                                hipStream_t s);
// Two-level sampling without a CDF array.  launch_sample_totals: blk (nb + 1
// doubles) <- FP64 chunk totals, blk[nb] = the total; qblk (nb + 1) <- the
// exclusive prefixes of the fixed-point chunk totals (weights round(|a|^2
// 2^62), valid while the total is below 4), qblk[nb] = their total; sums:
// sample_scan_words(nb) words of scan workspace.  launch_sample_chunks:
// out[s] = the first index whose inclusive fixed-point prefix exceeds
// floor(u[s] qblk[nb]).
struct SampleScratch {
    uint64_t *tgt;      // shots
    uint32_t *chunk;    // shots
    int64_t *bucket;    // shots
    uint32_t *counts;   // nb
    uint64_t *offsets;  // nb + 1
    uint64_t *sums;     // sample_scan_words(nb)
};
inline int64_t sample_scan_words(int64_t nb) { return (nb + 4095) / 4096 + 1; }
template <typename T>
hipError_t launch_sample_totals(const void *sv, int64_t dim, double *blk, uint64_t *qblk, uint64_t *sums,
                                hipStream_t s);
template <typename T>
hipError_t launch_sample_chunks(const void *sv, int64_t dim, const uint64_t *qblk, const double *u, int64_t shots,
                                SampleScratch w, int64_t *out, hipStream_t s);
template <typename T>
hipError_t launch_qubit_probs(const void *sv, int64_t dim, int bit, double *partials, int nblk, hipStream_t s);
template <typename T>
hipError_t launch_collapse(void *sv, int64_t dim, int bit, int outcome, double scale, hipStream_t s);
template <typename T>
hipError_t launch_pauli_expect(const void *sv, int n, const PauliItem *items, int n_items,
                               const uint64_t *zmask, double *partials, int nblk, hipStream_t s);
template <typename T>
hipError_t launch_dense_matvec(void *out, const void *in, int n, const double2 *H, hipStream_t s);
// Krylov bases: up to kMaxBasis state pointers per launch.
constexpr int kMaxBasis = 64;
constexpr int kInnerGroup = 8;
struct BasisPtrs {
    const void *p[kMaxBasis];
};
// partials[(2 j + {0, 1}) * nblk + b] for <vs_j|w>, j < m (then launch_sum_partials with 2 m values).
template <typename T>
hipError_t launch_inner_many(const void *w, const BasisPtrs &vs, int m, int64_t dim, double *partials, int nblk,
                             hipStream_t s);
// out_l <- sum_j coef[j * p + l] in_j (l < p, j < m <= kMaxBasis); outputs may alias inputs.
template <typename T>
hipError_t launch_combine(const BasisPtrs &in, const BasisPtrs &out, int m, int p, const double2 *coef, int64_t dim,
                          hipStream_t s);
// Reproducible pseudo-random amplitudes in [-1/2, 1/2)^2 from a 64-bit seed.
template <typename T>
hipError_t launch_fill_random(void *sv, int64_t dim, uint64_t seed, hipStream_t s);
// out[v] = sum_{b < nblk} partials[v * nblk + b], for v < nvals.
hipError_t launch_sum_partials(const double *partials, int nvals, int nblk, double *out, hipStream_t s);
// Max dynamic LDS bytes the stage kernel may use on this device.
size_t stage_lds_limit(int device);

}  // namespace desv
