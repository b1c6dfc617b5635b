// Device kernels of the dense statevector engine (gfx950, wave64).
//
// Every kernel works on interleaved complex amplitudes (float2 / double2).
// Reductions accumulate in FP64 and are written as per-block partials that a
// second, single-block pass sums in a fixed order, so results are
// reproducible run to run.
#pragma once

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

struct StageArgs {
    DevStage st;
    int n;
    const DevOp *ops;
    const void *mats;    // complex pool in the state precision
    double *partials;    // gradient partials [slot][n_wg]
    int n_wg;
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

// ---- launchers (engine_kernels.hip) --------------------------------------
template <typename T>
hipError_t launch_stage(int nv, void *sv0, void *sv1, const StageArgs &a, int grid, hipStream_t s);
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
template <typename T>
hipError_t launch_qubit_probs(const void *sv, int64_t dim, int bit, double *partials, int nblk, hipStream_t s);
template <typename T>
hipError_t launch_collapse(void *sv, int64_t dim, int bit, int outcome, double scale, hipStream_t s);
template <typename T>
hipError_t launch_pauli_expect(const void *sv, int n, const PauliItem *items, int n_items,
                               const uint64_t *zmask, double *partials, int nblk, hipStream_t s);
template <typename T>
hipError_t launch_pauli_matvec(void *out, const void *in, int n, const MatvecGroup *groups, int n_groups,
                               const uint64_t *zmask, const double2 *cc, hipStream_t s);
template <typename T>
hipError_t launch_dense_matvec(void *out, const void *in, int n, const double2 *H, hipStream_t s);
// out[v] = sum_{b < nblk} partials[v * nblk + b], for v < nvals.
hipError_t launch_sum_partials(const double *partials, int nvals, int nblk, double *out, hipStream_t s);
// Max dynamic LDS bytes the stage kernel may use on this device.
size_t stage_lds_limit(int device);

}  // namespace desv
