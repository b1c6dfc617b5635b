// C ABI of the engine (declared in include/desv.h): state handles, compiled
// programs, Pauli sums and the adjoint gradient driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/desv.h"
#include "desv_internal.h"
#include "kernels.h"

using namespace desv;

// Partials of desv_norm2_device, the entry point that allocates nothing: a
// file-scope device array (each GPU the library runs on holds its own copy
// until the process ends).  Host code never passes the array's name to a
// kernel: it uses the per-device address hipGetSymbolAddress returns for it.
constexpr int kScratchDoubles = 8192;  // 64 KiB: one FP64 partial per workgroup
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(e_ == hipErrorOutOfMemory ? DESV_ENOMEM : DESV_EDEVICE,                 \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                    \
    } while (0)

inline size_t amp_bytes(int dtype) { return dtype == DESV_C64 ? 8 : 16; }

// Device-side copy of a compiled plan (ops + matrices rounded to the state
// precision).
struct DevPlan {
    int device = -1;
    int dtype = -1;
    void *buf = nullptr;
    size_t cap = 0;
    DevOp *ops = nullptr;
    void *mats = nullptr;
    // compensated (complex64): mats holds the FP32 hi pool followed by the lo
    // pool, lo_off complex elements further on
    int comp = 0;
    int32_t lo_off = 0;
    void *ltab = nullptr;  // per-lane tables of 3-qubit MFMA ops (complex64), or null
    // register groups (complex64): device words and fragments, per-stage offsets
    int32_t *rwords = nullptr;
    float *rfrags = nullptr;
    std::vector<int32_t> rstage_word, rstage_frag, rstage_nq;  // plan.stages, then plan.fb_stages
    int rmode = 0;  // 1 register groups, 2 wave tiles
    int relabel_ok = 0;  // every relabelling stage has a wave program

    void release() {
        if (buf) (void)hipFree(buf);
        buf = nullptr;
        cap = 0;
    }
};

int upload_plan(const Plan &p, int dtype, int device, hipStream_t s, DevPlan &d, bool comp = false,
                int regs = 0) {
    comp = comp && dtype == DESV_C64;
    if (dtype != DESV_C64) regs = 0;
    if (comp && p.mats.size() > (size_t)0x3fffffff) return fail(DESV_EINVAL, "matrix pool too large");
    // matrices in the state precision (complex64: hi, then lo when compensated)
    const size_t nm = std::max<size_t>(p.mats.size(), 1) * (comp ? 2 : 1);
    std::vector<float> mf;
    std::vector<double> md;
    if (dtype == DESV_C64) {
        mf.assign(2 * nm, 0.f);
        round_mats_f32(p, mf.data());
        if (comp) round_mats_f32_lo(p, mf.data(), mf.data() + 2 * p.mats.size());
    } else {
        md.assign(2 * nm, 0.0);
        for (size_t i = 0; i < p.mats.size(); ++i) {
            md[2 * i] = p.mats[i].real();
            md[2 * i + 1] = p.mats[i].imag();
        }
    }
    // ops, with the per-lane tables of 3-qubit MFMA ops (complex64)
    std::vector<DevOp> ops = p.ops;
    std::vector<int32_t> lt;
    if (dtype == DESV_C64)
        build_lane_tables(p, mf.data(), comp ? mf.data() + 2 * p.mats.size() : nullptr, ops, lt);
    RegProgram rp;
    if (regs == 1) build_reg_program(p, comp, rp);
    if (regs == 2) build_wave_program(p, comp, rp);
    const size_t ob = std::max<size_t>(ops.size(), 1) * sizeof(DevOp);
    const size_t oba = (ob + 255) & ~size_t(255);
    const size_t mb = nm * amp_bytes(dtype);
    const size_t mba = (mb + 255) & ~size_t(255);
    const size_t lb = (lt.size() * sizeof(int32_t) + 255) & ~size_t(255);
    const size_t rwb = (rp.words.size() * sizeof(int32_t) + 255) & ~size_t(255);
    const size_t need = oba + mba + lb + rwb + rp.frags.size() * sizeof(float);
    HIP_TRY(hipStreamSynchronize(s));  // the previous upload may still be read
    if (d.cap < need || d.device != device) {
        d.release();
        HIP_TRY(hipMalloc(&d.buf, need));
        d.cap = need;
    }
    std::vector<unsigned char> host(need);
    std::memcpy(host.data(), ops.data(), ops.size() * sizeof(DevOp));
    if (dtype == DESV_C64)
        std::memcpy(host.data() + oba, mf.data(), mb);
    else
        std::memcpy(host.data() + oba, md.data(), mb);
    if (!lt.empty()) std::memcpy(host.data() + oba + mba, lt.data(), lt.size() * sizeof(int32_t));
    if (!rp.words.empty()) std::memcpy(host.data() + oba + mba + lb, rp.words.data(), rp.words.size() * sizeof(int32_t));
    if (!rp.frags.empty())
        std::memcpy(host.data() + oba + mba + lb + rwb, rp.frags.data(), rp.frags.size() * sizeof(float));
    HIP_TRY(hipMemcpyAsync(d.buf, host.data(), need, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    unsigned char *base = reinterpret_cast<unsigned char *>(d.buf);
    d.ops = reinterpret_cast<DevOp *>(d.buf);
    d.mats = base + oba;
    d.ltab = lt.empty() ? nullptr : base + oba + mba;
    d.rwords = rp.words.empty() ? nullptr : reinterpret_cast<int32_t *>(base + oba + mba + lb);
    d.rfrags = rp.words.empty() ? nullptr : reinterpret_cast<float *>(base + oba + mba + lb + rwb);
    d.rstage_word = rp.words.empty() ? std::vector<int32_t>() : rp.stage_word;
    d.rstage_frag = rp.stage_frag;
    d.rstage_nq.assign(rp.stage_word.size(), 0);
    for (size_t i = 0; i < rp.stage_word.size(); ++i)
        if (rp.stage_word[i] >= 0) d.rstage_nq[i] = rp.words[(size_t)rp.stage_word[i] + 1];
    d.rmode = rp.words.empty() ? 0 : regs;
    d.relabel_ok = regs == 2 && p.fb_from >= 0 && rp.relabel_ok && !rp.words.empty();
    d.device = device;
    d.dtype = dtype;
    d.comp = comp;
    d.lo_off = comp ? (int32_t)p.mats.size() : 0;
    return DESV_OK;
}

CompileOpts default_opts(int dtype, int device, int nv) {
    CompileOpts o;
    o.fuse_bits = 2;
    o.low_bits = 5;
    o.nv = nv;
    const size_t lim = stage_lds_limit(device);
    int m = 14;
    while (m > 6 && (size_t)nv * ((size_t)1 << m) * amp_bytes(dtype) > lim) --m;
    // 32 KB per vector keeps several workgroups resident per CU.
    const int pref = dtype == DESV_C64 ? 12 : 11;
    o.tile_bits = std::min(m, pref);
    o.per_gate = 0;
    o.use_mfma = 0;
    return o;
}

}  // namespace

int desv::set_last_error(int code, const std::string &msg) { return fail(code, msg); }

struct desv_state {
    int n = 0;
    int dtype = 0;
    int device = 0;
    int64_t dim = 0;
    size_t bytes = 0;
    void *sv = nullptr;
    void *aux[2] = {nullptr, nullptr};
    // Second state buffer for out-of-place streaming stages (read sv, write
    // alt, swap): allocated on first use unless double buffering is off.
    void *alt = nullptr;
    int double_buffer = 1;
    // complex64: gate tables of desv_run_ops / desv_apply_matrix are applied
    // with compensated (hi + lo) matrices
    int compensated = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    double *d_part = nullptr;
    size_t part_cap = 0;  // doubles
    double *d_out = nullptr;
    size_t out_cap = 0;   // doubles
    DevPlan scratch_plan;
    void *d_small = nullptr;  // term tables etc.
    size_t small_cap = 0;
};

struct desv_prog {
    int n = 0;
    std::vector<Gate> gates;
    int64_t n_logical = 0;
    bool compiled = false;
    CompileOpts opts;
    Plan plan;
    DevPlan dev;
};

namespace {

int check_handle(desv_handle h) {
    if (!h) return fail(DESV_EINVAL, "null handle");
    return DESV_OK;
}

int ensure_part(desv_handle h, size_t doubles) {
    if (h->part_cap >= doubles) return DESV_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->d_part) HIP_TRY(hipFree(h->d_part));
    h->d_part = nullptr;
    HIP_TRY(hipMalloc(&h->d_part, doubles * sizeof(double)));
    h->part_cap = doubles;
    return DESV_OK;
}

int ensure_out(desv_handle h, size_t doubles) {
    if (h->out_cap >= doubles) return DESV_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->d_out) HIP_TRY(hipFree(h->d_out));
    h->d_out = nullptr;
    HIP_TRY(hipMalloc(&h->d_out, doubles * sizeof(double)));
    h->out_cap = doubles;
    return DESV_OK;
}

int ensure_small(desv_handle h, size_t bytes) {
    if (h->small_cap >= bytes) return DESV_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->d_small) HIP_TRY(hipFree(h->d_small));
    h->d_small = nullptr;
    HIP_TRY(hipMalloc(&h->d_small, bytes));
    h->small_cap = bytes;
    return DESV_OK;
}

int ensure_aux(desv_handle h, int i) {
    if (h->aux[i]) return DESV_OK;
    HIP_TRY(hipMalloc(&h->aux[i], h->bytes));
    return DESV_OK;
}

// Workgroups of a grid-stride FP64 reduction: one per 4 x 256 amplitudes, at
// most 8192 (32 waves per CU: enough loads in flight to stream a large state
// from HBM).
int reduce_blocks(int64_t dim) {
    int64_t b = (dim + kThreads * 4 - 1) / (kThreads * 4);
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 8192));
}

// Read `nvals` sums of partials back to the host (synchronous).
int sum_to_host(desv_handle h, int nvals, int nblk, double *out) {
    int rc = ensure_out(h, nvals);
    if (rc) return rc;
    HIP_TRY(launch_sum_partials(h->d_part, nvals, nblk, h->d_out, h->stream));
    HIP_TRY(hipMemcpyAsync(out, h->d_out, nvals * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return DESV_OK;
}

// A stage that is exactly one dense 1-/2-qubit or diagonal (k <= 2) op with
// no write-back permutation runs on the streaming direct kernel.
bool direct_op(const Plan &plan, const DevStage &st, DirectOp &d) {
    if (st.op_end - st.op_begin != 1 || st.permuted) return false;
    const DevOp &op = plan.ops[st.op_begin];
    if (op.kind != OPK_DENSE1 && op.kind != OPK_DENSE2 && op.kind != OPK_DIAG) return false;
    if (op.k < 1 || op.k > 2) return false;
    int gb[64], c = 0;
    for (int b = 0; b < plan.n; ++b)
        if (st.tile_mask >> b & 1) gb[c++] = b;
    std::memset(&d, 0, sizeof(d));
    d.k = op.k;
    d.diag = op.kind == OPK_DIAG;
    for (int j = 0; j < op.k; ++j) d.pos[j] = op.pos[j] >= 0 ? gb[op.pos[j]] : -1 - op.pos[j];
    d.cmask = op.cglob;
    for (int j = 0; j < 32; ++j)
        if (op.cloc >> j & 1) d.cmask |= 1ull << gb[j];
    return true;
}

// The second buffer for out-of-place streaming stages, or nullptr when double
// buffering is off or the allocation fails (then the stages run in place).
void *ensure_alt(desv_handle h) {
    if (!h->double_buffer) return nullptr;
    if (!h->alt && hipMalloc(&h->alt, h->bytes) != hipSuccess) {
        (void)hipGetLastError();
        h->alt = nullptr;
        h->double_buffer = 0;
    }
    return h->alt;
}

// Launch every stage of a plan (one tile per workgroup for LDS stages).
// Gradient probes write partials[slot * pstride + tile], pstride = 2^(n - m).
// When sv0 is the handle's own state, a
// streaming stage may run out of place into the second buffer and swap it
// with the state (the handle's state pointer changes; callers re-read it).
// `evs` (optional, stages + 1 events) are recorded before every stage and
// after the last one.
int run_plan_on(desv_handle h, const Plan &plan, const DevPlan &dp, int nv, void *sv0, void *sv1,
                int64_t pstride, hipEvent_t *evs = nullptr, size_t *n_run = nullptr) {
    const bool own_state = nv == 1 && sv0 == h->sv;
    // the stages to run (with their register-program index): a relabelled
    // tail needs the second buffer, else its in-place form runs
    const size_t S = plan.stages.size();
    const bool relabel = plan.fb_from >= 0 && dp.relabel_ok && dp.rmode == 2 && own_state && ensure_alt(h);
    std::vector<std::pair<const DevStage *, size_t>> list;
    const size_t head = plan.fb_from < 0 || relabel ? S : (size_t)plan.fb_from;
    for (size_t i = 0; i < head; ++i) list.emplace_back(&plan.stages[i], i);
    if (plan.fb_from >= 0 && !relabel)
        for (size_t k = 0; k < plan.fb_stages.size(); ++k) list.emplace_back(&plan.fb_stages[k], S + k);
    if (n_run) *n_run = list.size();
    size_t si = 0;
    for (const auto &entry : list) {
        const DevStage &st = *entry.first;
        const size_t ri = entry.second;
        const bool relabelled = relabel && ri < S && !plan.relabel.empty() && plan.relabel[ri].store_mask;
        if (evs) HIP_TRY(hipEventRecord(evs[si], h->stream));
        ++si;
        DirectOp dop;
        // compensated programs run every stage on the stage kernel
        if (nv == 1 && !dp.comp && direct_op(plan, st, dop)) {
            const size_t eb = h->dtype == DESV_C64 ? 8 : 16;
            const void *M = reinterpret_cast<const char *>(dp.mats) + (size_t)plan.ops[st.op_begin].mat * eb;
            void *out = own_state ? ensure_alt(h) : nullptr;
            if (!out) out = sv0;
            hipError_t e = h->dtype == DESV_C64 ? launch_direct<float>(sv0, out, plan.n, dop, M, h->stream)
                                                : launch_direct<double>(sv0, out, plan.n, dop, M, h->stream);
            if (e != hipSuccess) return fail(DESV_EDEVICE, std::string("direct launch: ") + hipGetErrorString(e));
            if (out != sv0) {
                h->alt = sv0;
                h->sv = out;
                sv0 = out;
            }
            continue;
        }
        StageTask t;
        std::memset(&t, 0, sizeof(t));
        t.st = st;
        t.n = plan.n;
        t.sv0 = sv0;
        t.sv1 = sv1;
        t.ops = dp.ops;
        t.mats = dp.mats;
        t.lo_off = dp.lo_off;
        t.ltab = dp.ltab;
        t.partials = h->d_part;
        t.pstride = pstride;
        t.wg_begin = 0;
        int variant = st.heavy;
        if (nv == 1 && ri < dp.rstage_word.size() && dp.rstage_word[ri] >= 0) {
            // register-resident tile
            t.rprog = dp.rwords + dp.rstage_word[ri];
            t.rfrag = dp.rfrags + (size_t)dp.rstage_frag[ri] * 256;
            if (dp.rmode == 2) {
                if (relabelled) t.sv1 = h->alt;
                const hipError_t e = launch_wave_stage(st.m, dp.comp != 0, dp.rstage_nq[ri], t, h->stream);
                if (e != hipSuccess) return fail(DESV_EDEVICE, std::string("wave stage launch: ") + hipGetErrorString(e));
                if (relabelled) {
                    h->alt = sv0;
                    h->sv = t.sv1;
                    sv0 = h->sv;
                }
                continue;
            }
            variant = 4;
        }
        if (relabelled) return fail(DESV_EINVAL, "internal: relabelling stage without a wave program");
        hipError_t e = h->dtype == DESV_C64
                           ? launch_stage<float>(nv, variant, dp.comp != 0, nullptr, nullptr, t, st.n_tiles, st.m,
                                                 h->stream)
                           : launch_stage<double>(nv, st.heavy, false, nullptr, nullptr, t, st.n_tiles, st.m, h->stream);
        if (e != hipSuccess) return fail(DESV_EDEVICE, std::string("stage launch: ") + hipGetErrorString(e));
    }
    if (evs) HIP_TRY(hipEventRecord(evs[si], h->stream));
    return DESV_OK;
}

int decode_ops(const double *ops, int64_t n_ops, int n, std::vector<Gate> &gates) {
    std::string err;
    for (int64_t i = 0; i < n_ops; ++i) {
        size_t before = gates.size();
        if (!decode_row(ops + 4 * i, n, gates, err)) return fail(DESV_EINVAL, err);
        for (size_t j = before; j < gates.size(); ++j) gates[j].src = (int)i;
    }
    return DESV_OK;
}

int apply_gates(desv_handle h, const std::vector<Gate> &gates) {
    if (gates.empty()) return DESV_OK;
    CompileOpts o = default_opts(h->dtype, h->device, 1);
    Plan plan;
    std::string err;
    if (!build_plan(gates, h->n, o, plan, err)) return fail(DESV_EINVAL, err);
    int rc = upload_plan(plan, h->dtype, h->device, h->stream, h->scratch_plan, h->compensated != 0);
    if (rc) return rc;
    return run_plan_on(h, plan, h->scratch_plan, 1, h->sv, nullptr, 0);
}

// Pauli-sum matvec into `out` (device buffer) from `in`.
// The tiled Pauli-sum kernel on `in`: lambda = H in into `out` (may be null)
// and, if `epart` is set, one partial Re<in|H in> per workgroup (`*grid_out`
// workgroups).
int pauli_tiles_raw(desv_handle h, void *out, const void *in, const uint64_t *xmask, const uint64_t *zmask,
                    const double *cre, const double *cim, int64_t T, double *epart, int64_t *grid_out) {
    // group terms by flip mask; fold i^{#Y} into the coefficient
    std::vector<int64_t> order(T);
    for (int64_t i = 0; i < T; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return xmask[a] < xmask[b]; });
    std::vector<MatvecGroup> groups;
    std::vector<uint64_t> z(T);
    std::vector<double2> cc(T);
    for (int64_t r = 0; r < T; ++r) {
        const int64_t t = order[r];
        z[r] = zmask[t];
        const int ny = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
        double re = cre[t], im = cim ? cim[t] : 0.0, a = re, b = im;
        // multiply by i^ny
        for (int k = 0; k < ny; ++k) {
            double t2 = a;
            a = -b;
            b = t2;
        }
        cc[r] = make_double2(a, b);
        if (groups.empty() || groups.back().x != xmask[t]) groups.push_back({xmask[t], (int32_t)r, (int32_t)r + 1});
        else groups.back().end = (int32_t)r + 1;
    }
    const size_t gb = groups.size() * sizeof(MatvecGroup), zb = T * sizeof(uint64_t), cb = T * sizeof(double2);
    const size_t ga = (gb + 255) & ~size_t(255), za = (zb + 255) & ~size_t(255), ca = (cb + 255) & ~size_t(255);
    int rc = ensure_small(h, ga + za + ca + sizeof(MatvecTask) + 256);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    unsigned char *base = reinterpret_cast<unsigned char *>(h->d_small);
    // one task of the tiled matvec kernel (no psi copy)
    MatvecTask tk;
    std::memset(&tk, 0, sizeof(tk));
    tk.n = h->n;
    tk.ngroups = (int32_t)groups.size();
    tk.psi = in;
    tk.lam = out;
    tk.epart = epart;
    tk.groups = base;
    tk.z = reinterpret_cast<const uint64_t *>(base + ga);
    tk.cc = reinterpret_cast<const double2 *>(base + ga + za);
    tk.chunk = 1ll << matvec_tile_bits(h->n);
    const int64_t grid = (1ll << h->n) / tk.chunk;
    if (grid_out) *grid_out = grid;
    std::vector<unsigned char> host(ga + za + ca + sizeof(MatvecTask));
    std::memcpy(host.data(), groups.data(), gb);
    std::memcpy(host.data() + ga, z.data(), zb);
    std::memcpy(host.data() + ga + za, cc.data(), cb);
    std::memcpy(host.data() + ga + za + ca, &tk, sizeof(tk));
    HIP_TRY(hipMemcpyAsync(h->d_small, host.data(), host.size(), hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    const MatvecTask *dtk = reinterpret_cast<const MatvecTask *>(base + ga + za + ca);
    hipError_t e = h->dtype == DESV_C64 ? launch_matvec_tasks<float>(dtk, 1, grid, h->stream)
                                        : launch_matvec_tasks<double>(dtk, 1, grid, h->stream);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    return DESV_OK;
}

int pauli_matvec_raw(desv_handle h, void *out, const void *in, const uint64_t *xmask, const uint64_t *zmask,
                     const double *cre, const double *cim, int64_t T) {
    return pauli_tiles_raw(h, out, in, xmask, zmask, cre, cim, T, nullptr, nullptr);
}

int check_masks(desv_handle h, const uint64_t *xmask, const uint64_t *zmask, int64_t T) {
    const uint64_t all = (h->n >= 64) ? ~0ull : ((1ull << h->n) - 1);
    for (int64_t t = 0; t < T; ++t)
        if ((xmask[t] | zmask[t]) & ~all) return fail(DESV_EINVAL, "Pauli term references a qubit out of range");
    return DESV_OK;
}

int dense_matvec_raw(desv_handle h, void *out, const void *in, const double *H) {
    if (h->n > 14) return fail(DESV_EINVAL, "dense Hamiltonians are limited to n <= 14");
    const size_t bytes = (size_t)h->dim * h->dim * 16;
    int rc = ensure_small(h, bytes);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpyAsync(h->d_small, H, bytes, hipMemcpyHostToDevice, h->stream));
    hipError_t e = h->dtype == DESV_C64
                       ? launch_dense_matvec<float>(out, in, h->n, (const double2 *)h->d_small, h->stream)
                       : launch_dense_matvec<double>(out, in, h->n, (const double2 *)h->d_small, h->stream);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return DESV_OK;
}

int inner_raw(desv_handle h, const void *a, const void *b, double *re, double *im) {
    const int nblk = reduce_blocks(h->dim);
    int rc = ensure_part(h, 2 * (size_t)nblk);
    if (rc) return rc;
    hipError_t e = h->dtype == DESV_C64 ? launch_inner<float>(a, b, h->dim, h->d_part, nblk, h->stream)
                                        : launch_inner<double>(a, b, h->dim, h->d_part, nblk, h->stream);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    double out[2];
    rc = sum_to_host(h, 2, nblk, out);
    if (rc) return rc;
    if (re) *re = out[0];
    if (im) *im = out[1];
    return DESV_OK;
}

}  // namespace

extern "C" {

const char *desv_last_error(void) { return g_err.c_str(); }

int desv_version(void) { return 2; }

int desv_device_count(int *out) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    if (out) *out = c;
    return DESV_OK;
}

int desv_create(int n, int dtype, int device, desv_handle *out) {
    if (!out) return fail(DESV_EINVAL, "null output pointer");
    *out = nullptr;
    if (n < 1 || n > 34) return fail(DESV_EINVAL, "n_qubits must be in [1, 34], got " + std::to_string(n));
    if (dtype != DESV_C64 && dtype != DESV_C128) return fail(DESV_EINVAL, "dtype must be DESV_C64 or DESV_C128");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(DESV_EDEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) return fail(DESV_EINVAL, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    std::unique_ptr<desv_state> h(new (std::nothrow) desv_state());
    if (!h) return fail(DESV_ENOMEM, "host allocation failed");
    h->n = n;
    h->dtype = dtype;
    h->device = device;
    h->dim = 1ll << n;
    h->bytes = (size_t)h->dim * amp_bytes(dtype);
    HIP_TRY(hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking));
    h->stream = h->own;
    hipError_t e = hipMalloc(&h->sv, h->bytes);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(h->own);
        return fail(DESV_ENOMEM, "cannot allocate " + std::to_string(h->bytes) + " bytes of device memory");
    }
    (void)stage_lds_limit(device);
    e = dtype == DESV_C64 ? launch_fill_basis<float>(h->sv, h->dim, 0, h->stream)
                          : launch_fill_basis<double>(h->sv, h->dim, 0, h->stream);
    if (e != hipSuccess) {
        (void)hipFree(h->sv);
        (void)hipStreamDestroy(h->own);
        return fail(DESV_EDEVICE, hipGetErrorString(e));
    }
    *out = h.release();
    return DESV_OK;
}

int desv_destroy(desv_handle h) {
    if (!h) return DESV_OK;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    if (h->sv) (void)hipFree(h->sv);
    if (h->alt) (void)hipFree(h->alt);
    for (void *a : h->aux)
        if (a) (void)hipFree(a);
    if (h->d_part) (void)hipFree(h->d_part);
    if (h->d_out) (void)hipFree(h->d_out);
    if (h->d_small) (void)hipFree(h->d_small);
    h->scratch_plan.release();
    if (h->own) (void)hipStreamDestroy(h->own);
    delete h;
    return DESV_OK;
}

int desv_set_stream(desv_handle h, void *stream) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->stream = stream ? (hipStream_t)stream : h->own;
    return DESV_OK;
}

int desv_set_double_buffer(desv_handle h, int enable) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->double_buffer = enable ? 1 : 0;
    if (!enable && h->alt) {
        HIP_TRY(hipFree(h->alt));
        h->alt = nullptr;
    }
    return DESV_OK;
}

int desv_state_device_ptr(desv_handle h, void **out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(DESV_EINVAL, "null output pointer");
    *out = h->sv;
    return DESV_OK;
}

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[64] = {};

// Device address of `scratch` on `device`, looked up once per device with
// that device current (the caller has made it current).
int scratch_on(int device, double **out) {
    if (device < 0 || device >= 64) return fail(DESV_EINVAL, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(DESV_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return DESV_OK;
}
}  // namespace

int desv_norm2_device(desv_handle h, double *out_dev) {
    if (int rc = check_handle(h)) return rc;
    if (!out_dev) return fail(DESV_EINVAL, "null output pointer");
    HIP_TRY(hipSetDevice(h->device));
    const int nblk = reduce_blocks(h->dim);
    if (nblk > kScratchDoubles)
        return fail(DESV_EINVAL, "the reduction needs " + std::to_string(nblk) + " partials; the scratch holds " +
                                     std::to_string(kScratchDoubles));
    double *part = nullptr;
    if (int rc = scratch_on(h->device, &part)) return rc;
    // every workgroup writes its own partial, so nothing needs clearing
    HIP_TRY(h->dtype == DESV_C64 ? launch_norm2<float>(h->sv, h->dim, part, nblk, h->stream)
                                 : launch_norm2<double>(h->sv, h->dim, part, nblk, h->stream));
    HIP_TRY(launch_sum_partials(part, 1, nblk, out_dev, h->stream));
    return DESV_OK;
}

int desv_set_compensated(desv_handle h, int enable) {
    if (int rc = check_handle(h)) return rc;
    h->compensated = enable && h->dtype == DESV_C64 ? 1 : 0;
    return DESV_OK;
}

int desv_synchronize(desv_handle h) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return DESV_OK;
}

int desv_reset(desv_handle h) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(h->dtype == DESV_C64 ? launch_fill_basis<float>(h->sv, h->dim, 0, h->stream)
                                 : launch_fill_basis<double>(h->sv, h->dim, 0, h->stream));
    return DESV_OK;
}

int desv_set_state(desv_handle h, const void *host) {
    if (int rc = check_handle(h)) return rc;
    if (!host) return fail(DESV_EINVAL, "null host buffer");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipMemcpyAsync(h->sv, host, h->bytes, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return DESV_OK;
}

int desv_get_state(desv_handle h, void *host) {
    if (int rc = check_handle(h)) return rc;
    if (!host) return fail(DESV_EINVAL, "null host buffer");
    HIP_TRY(hipSetDevice(h->device));
    HIP_TRY(hipMemcpyAsync(host, h->sv, h->bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return DESV_OK;
}

int desv_copy(desv_handle dst, desv_handle src) {
    if (int rc = check_handle(dst)) return rc;
    if (int rc = check_handle(src)) return rc;
    if (dst->n != src->n) return fail(DESV_EINVAL, "copy between different shapes");
    if (dst->device != src->device) return fail(DESV_EINVAL, "copy between different devices");
    HIP_TRY(hipSetDevice(dst->device));
    HIP_TRY(hipStreamSynchronize(src->stream));
    if (dst->dtype == src->dtype && dst->bytes % 16 == 0) {
        HIP_TRY(launch_copy16(dst->sv, src->sv, dst->bytes, dst->stream));
    } else if (dst->dtype == src->dtype) {
        HIP_TRY(hipMemcpyAsync(dst->sv, src->sv, dst->bytes, hipMemcpyDeviceToDevice, dst->stream));
    } else {
        const hipError_t e = dst->dtype == DESV_C64
                                 ? launch_convert<float, double>(dst->sv, src->sv, dst->dim, dst->stream)
                                 : launch_convert<double, float>(dst->sv, src->sv, dst->dim, dst->stream);
        HIP_TRY(e);
    }
    return DESV_OK;
}

int desv_scale(desv_handle h, double re, double im) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(h->dtype == DESV_C64 ? launch_scale<float>(h->sv, h->dim, re, im, h->stream)
                                 : launch_scale<double>(h->sv, h->dim, re, im, h->stream));
    return DESV_OK;
}

int desv_norm2(desv_handle h, double *out) {
    if (int rc = check_handle(h)) return rc;
    const int nblk = reduce_blocks(h->dim);
    int rc = ensure_part(h, nblk);
    if (rc) return rc;
    HIP_TRY(h->dtype == DESV_C64 ? launch_norm2<float>(h->sv, h->dim, h->d_part, nblk, h->stream)
                                 : launch_norm2<double>(h->sv, h->dim, h->d_part, nblk, h->stream));
    return sum_to_host(h, 1, nblk, out);
}

int desv_inner(desv_handle a, desv_handle b, double *re, double *im) {
    if (int rc = check_handle(a)) return rc;
    if (int rc = check_handle(b)) return rc;
    if (a->n != b->n || a->dtype != b->dtype) return fail(DESV_EINVAL, "inner product of different shapes");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return inner_raw(a, a->sv, b->sv, re, im);
}

int desv_probabilities(desv_handle h, double *out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(DESV_EINVAL, "null output buffer");
    double *d = nullptr;
    HIP_TRY(hipMalloc(&d, h->dim * sizeof(double)));
    hipError_t e = h->dtype == DESV_C64 ? launch_probs<float>(h->sv, h->dim, d, h->stream)
                                        : launch_probs<double>(h->sv, h->dim, d, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, h->dim * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    return DESV_OK;
}

int desv_sample(desv_handle h, const double *uniforms, int64_t n_shots, int64_t *out) {
    if (int rc = check_handle(h)) return rc;
    if (n_shots < 1) return fail(DESV_EINVAL, "n_shots must be >= 1, got " + std::to_string(n_shots));
    if (!uniforms || !out) return fail(DESV_EINVAL, "null sample buffers");
    HIP_TRY(hipSetDevice(h->device));
    const int64_t nb = cdf_blocks(h->dim);
    if (int rc = ensure_part(h, 2 * ((size_t)nb + 1) + (size_t)sample_scan_words(nb))) return rc;
    {
        // Two-level sampling (chunk totals, then the chunks that hold shots):
        // the fixed-point weights use the fixed scale 2^62, which keeps the
        // 64-bit sums exact for a total probability below 4 (and within a
        // factor 16 of the CDF path's quantisation above 1/16).
        double *blk = h->d_part;
        uint64_t *qblk = reinterpret_cast<uint64_t *>(blk + nb + 1);
        uint64_t *sums = qblk + nb + 1;
        hipError_t e = h->dtype == DESV_C64 ? launch_sample_totals<float>(h->sv, h->dim, blk, qblk, sums, h->stream)
                                            : launch_sample_totals<double>(h->sv, h->dim, blk, qblk, sums, h->stream);
        if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
        double total = 0;
        HIP_TRY(hipMemcpyAsync(&total, blk + nb, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (!(total > 0) || !std::isfinite(total))
            return fail(DESV_EINVAL, "statevector has zero or non-finite total probability");
        if (total >= 0.0625 && total < 3.5) {
            const size_t S = (size_t)n_shots, B = (size_t)nb;
            auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
            const size_t o_u = 0, o_out = o_u + al(S * 8), o_tgt = o_out + al(S * 8), o_chunk = o_tgt + al(S * 8),
                         o_bucket = o_chunk + al(S * 4), o_counts = o_bucket + al(S * 8),
                         o_offsets = o_counts + al(B * 4), bytes = o_offsets + al((B + 1) * 8);
            unsigned char *w = nullptr;
            e = hipMalloc(&w, bytes);
            if (e != hipSuccess) return fail(DESV_ENOMEM, std::string("sample scratch: ") + hipGetErrorString(e));
            SampleScratch sc{reinterpret_cast<uint64_t *>(w + o_tgt), reinterpret_cast<uint32_t *>(w + o_chunk),
                             reinterpret_cast<int64_t *>(w + o_bucket), reinterpret_cast<uint32_t *>(w + o_counts),
                             reinterpret_cast<uint64_t *>(w + o_offsets), sums};
            double *du = reinterpret_cast<double *>(w + o_u);
            int64_t *dout = reinterpret_cast<int64_t *>(w + o_out);
            e = hipMemcpyAsync(du, uniforms, S * sizeof(double), hipMemcpyHostToDevice, h->stream);
            if (e == hipSuccess)
                e = h->dtype == DESV_C64
                        ? launch_sample_chunks<float>(h->sv, h->dim, qblk, du, n_shots, sc, dout, h->stream)
                        : launch_sample_chunks<double>(h->sv, h->dim, qblk, du, n_shots, sc, dout, h->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(out, dout, S * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            (void)hipFree(w);
            if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
            return DESV_OK;
        }
    }
    // far from normalised: the CDF path, whose fixed-point scale follows the total
    if (int rc = ensure_aux(h, 1)) return rc;  // the CDF (2^n doubles fit in a state buffer)
    double *cdf = reinterpret_cast<double *>(h->aux[1]);
    hipError_t e = h->dtype == DESV_C64 ? launch_cdf<float>(h->sv, h->dim, cdf, h->d_part, h->stream)
                                        : launch_cdf<double>(h->sv, h->dim, cdf, h->d_part, h->stream);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    double total = 0;
    HIP_TRY(hipMemcpyAsync(&total, h->d_part + nb, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!(total > 0) || !std::isfinite(total))
        return fail(DESV_EINVAL, "statevector has zero or non-finite total probability");
    double *du = nullptr;
    int64_t *dout = nullptr;
    e = hipMalloc(&du, (size_t)n_shots * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dout, (size_t)n_shots * sizeof(int64_t));
    if (e == hipSuccess)
        e = hipMemcpyAsync(du, uniforms, (size_t)n_shots * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = launch_sample_search(cdf, h->dim, du, n_shots, dout, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, dout, (size_t)n_shots * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (du) (void)hipFree(du);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    return DESV_OK;
}

int desv_qubit_probs(desv_handle h, int qubit, double *p0, double *p1) {
    if (int rc = check_handle(h)) return rc;
    if (qubit < 0 || qubit >= h->n) return fail(DESV_EINVAL, "Qubit " + std::to_string(qubit) + " out of range");
    const int nblk = reduce_blocks(h->dim);
    int rc = ensure_part(h, 2 * (size_t)nblk);
    if (rc) return rc;
    const int bit = h->n - 1 - qubit;
    HIP_TRY(h->dtype == DESV_C64 ? launch_qubit_probs<float>(h->sv, h->dim, bit, h->d_part, nblk, h->stream)
                                 : launch_qubit_probs<double>(h->sv, h->dim, bit, h->d_part, nblk, h->stream));
    double out[2];
    rc = sum_to_host(h, 2, nblk, out);
    if (rc) return rc;
    if (p0) *p0 = out[0];
    if (p1) *p1 = out[1];
    return DESV_OK;
}

int desv_apply1q_probs(desv_handle h, int qubit, const double *m, int next_qubit, double *p0, double *p1) {
    if (int rc = check_handle(h)) return rc;
    if (qubit < 0 || qubit >= h->n) return fail(DESV_EINVAL, "Qubit " + std::to_string(qubit) + " out of range");
    if (next_qubit >= h->n) return fail(DESV_EINVAL, "Qubit " + std::to_string(next_qubit) + " out of range");
    if (!m) return fail(DESV_EINVAL, "null matrix");
    if (next_qubit >= 0 && (!p0 || !p1)) return fail(DESV_EINVAL, "null probability outputs");
    HIP_TRY(hipSetDevice(h->device));
    // one streaming pass: enough workgroups to keep the pairs' loads in flight
    const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(h->dim / (2 * kThreads * 8), 8192));
    if (int rc = ensure_part(h, 2 * (size_t)nblk)) return rc;
    const int bit = h->n - 1 - qubit, pbit = next_qubit >= 0 ? h->n - 1 - next_qubit : -1;
    HIP_TRY(h->dtype == DESV_C64 ? launch_apply1q_probs<float>(h->sv, h->dim, bit, m, pbit, h->d_part, nblk, h->stream)
                                 : launch_apply1q_probs<double>(h->sv, h->dim, bit, m, pbit, h->d_part, nblk, h->stream));
    if (next_qubit < 0) return DESV_OK;
    double out[2];
    if (int rc = sum_to_host(h, 2, nblk, out)) return rc;
    *p0 = out[0];
    *p1 = out[1];
    return DESV_OK;
}

int desv_apply_monomial_layer(desv_handle h, int k, const int *qubits, const double *mats, int n_pop,
                              const int *pop_qubits, double *pops) {
    if (int rc = check_handle(h)) return rc;
    const int n = h->n;
    if (k < 0 || k > n || (k > 0 && (!qubits || !mats))) return fail(DESV_EINVAL, "bad operator list");
    if (pops && (n_pop < 0 || n_pop > std::min(n, kMonoMaxPops) || (n_pop > 0 && !pop_qubits)))
        return fail(DESV_EINVAL, "population qubits: at most " + std::to_string(kMonoMaxPops));
    if (k == 0 && !pops) return DESV_OK;
    const int pvb = h->dtype == DESV_C64 ? 1 : 0;
    uint64_t F = 0, used = 0;
    struct Fac {
        int bit;
        double2 c[2];
    };
    std::vector<Fac> fac;
    for (int i = 0; i < k; ++i) {
        const int q = qubits[i];
        if (q < 0 || q >= n) return fail(DESV_EINVAL, "Qubit " + std::to_string(q) + " out of range");
        const int bit = n - 1 - q;
        if (used >> bit & 1) return fail(DESV_EINVAL, "qubits of a layer must differ");
        used |= 1ull << bit;
        const double *m = mats + 8 * i;
        int flip = -1;
        Fac f{bit, {make_double2(0, 0), make_double2(0, 0)}};
        for (int r = 0; r < 2; ++r) {
            int col = -1;
            for (int c = 0; c < 2; ++c)
                if (m[4 * r + 2 * c] != 0.0 || m[4 * r + 2 * c + 1] != 0.0) {
                    if (col >= 0) return fail(DESV_EINVAL, "operator on qubit " + std::to_string(q) +
                                                               " has two nonzero entries in a row");
                    col = c;
                }
            if (col < 0) continue;
            if (flip >= 0 && flip != (r ^ col))
                return fail(DESV_EINVAL, "operator on qubit " + std::to_string(q) + " maps both rows to one column");
            flip = r ^ col;
            f.c[r] = make_double2(m[4 * r + 2 * col], m[4 * r + 2 * col + 1]);
        }
        if (flip > 0) F |= 1ull << bit;
        fac.push_back(f);
    }
    std::vector<int> pbit(pops ? n_pop : 0);
    uint64_t pused = 0;
    for (size_t j = 0; j < pbit.size(); ++j) {
        const int q = pop_qubits[j];
        if (q < 0 || q >= n) return fail(DESV_EINVAL, "Qubit " + std::to_string(q) + " out of range");
        pbit[j] = n - 1 - q;
        if (pused >> pbit[j] & 1) return fail(DESV_EINVAL, "population qubits must differ");
        pused |= 1ull << pbit[j];
    }
    HIP_TRY(hipSetDevice(h->device));
    MonoTask t{};
    const int nvb = n - pvb;
    t.nvec = h->dim >> pvb;
    t.f0 = pvb ? (int)(F & 1) : 0;
    t.fv = (int64_t)(F >> pvb);
    t.hb = t.fv ? 63 - __builtin_clzll((uint64_t)t.fv) : -1;
    t.lb = (t.fv && t.hb < 6) ? 7 : 6;
    t.lamp = t.lb + pvb;
    t.store = k > 0;
    std::vector<double2> tab;
    for (const Fac &f : fac) {
        if (f.bit < t.lamp) {
            t.lbit[t.nlow] = f.bit;
            t.lc[t.nlow][0] = f.c[0];
            t.lc[t.nlow][1] = f.c[1];
            ++t.nlow;
        } else {
            t.ntab = std::max(t.ntab, (f.bit - t.lamp) / 8 + 1);
        }
    }
    if (t.ntab > kMonoTables) return fail(DESV_EINVAL, "too many qubits for a layer");
    if (t.ntab) {
        tab.assign((size_t)t.ntab * 256, make_double2(1.0, 0.0));
        for (const Fac &f : fac) {
            if (f.bit < t.lamp) continue;
            const int kk = (f.bit - t.lamp) / 8, sh = (f.bit - t.lamp) % 8;
            for (int by = 0; by < 256; ++by) {
                double2 &e = tab[(size_t)kk * 256 + by];
                const double2 c = f.c[(by >> sh) & 1];
                e = make_double2(e.x * c.x - e.y * c.y, e.x * c.y + e.y * c.x);
            }
        }
        if (int rc = ensure_small(h, tab.size() * sizeof(double2))) return rc;
        HIP_TRY(hipMemcpyAsync(h->d_small, tab.data(), tab.size() * sizeof(double2), hipMemcpyHostToDevice,
                               h->stream));
        t.tab = (const double2 *)h->d_small;
    }
    const int hbv = (t.fv && t.hb >= t.lb) ? t.hb : -1;
    t.hbit_amp = hbv >= 0 ? 1ull << (hbv + pvb) : 0;
    for (int b : pbit) {
        if (b < t.lamp || b - pvb == hbv) t.plmask |= 1ull << b;
        else t.phmask |= 1ull << (b - pvb);
    }
    for (int b = t.lb; b < nvb; ++b)
        if (b != hbv && !(t.phmask >> b & 1)) t.rmask |= 1ull << b;
    t.rstep1 = t.rmask & (~t.rmask + 1);
    t.nruns = 1ll << __builtin_popcountll(t.rmask);
    t.npl = __builtin_popcountll(t.plmask);
    for (int i = 0, r = 0; i < 64; ++i)
        if (t.phmask >> i & 1) {
            if ((uint64_t)t.fv >> i & 1) t.fph |= 1u << r;
            ++r;
        }
    const int ns = t.fv ? 2 : 1;
    const int gy = 1 << __builtin_popcountll(t.phmask);
#ifndef DESV_MONO_WGS
#define DESV_MONO_WGS 8192
#endif
    const int64_t want = std::max<int64_t>(1, DESV_MONO_WGS / gy);
    const int64_t per_wg = 16;  // runs per workgroup iteration (4 waves x 4 runs)
    int64_t gx = std::min<int64_t>(want, std::max<int64_t>(1, (t.nruns + per_wg - 1) / per_wg));
    if (pops) gx = std::min<int64_t>(gx, std::max<int64_t>(1, (1ll << 22) / ((int64_t)ns << n_pop)));
    if (pops) {
        if (int rc = ensure_part(h, ((size_t)1 << n_pop) * gx * ns)) return rc;
        t.partials = h->d_part;
    }
    HIP_TRY(h->dtype == DESV_C64 ? launch_mono_layer<float>(h->sv, t, (int)gx, gy, h->stream)
                                 : launch_mono_layer<double>(h->sv, t, (int)gx, gy, h->stream));
    if (!pops) {
        if (t.ntab) HIP_TRY(hipStreamSynchronize(h->stream));  // the host table outlives the copy
        return DESV_OK;
    }
    std::vector<double> comp((size_t)1 << n_pop);
    if (int rc = sum_to_host(h, 1 << n_pop, (int)gx * ns, comp.data())) return rc;
    // composite bin: plmask bits (ascending) below, phmask bits (ascending) above
    std::vector<int> at(pbit.size());
    for (size_t j = 0; j < pbit.size(); ++j) {
        const int b = pbit[j];
        if (t.plmask >> b & 1) at[j] = __builtin_popcountll(t.plmask & ((1ull << b) - 1));
        else at[j] = t.npl + __builtin_popcountll(t.phmask & ((1ull << (b - pvb)) - 1));
    }
    for (size_t c = 0; c < comp.size(); ++c) {
        size_t idx = 0;
        for (size_t j = 0; j < pbit.size(); ++j) idx |= (size_t)((c >> at[j]) & 1) << j;
        pops[idx] = comp[c];
    }
    return DESV_OK;
}

int desv_collapse(desv_handle h, int qubit, int outcome, double scale) {
    if (int rc = check_handle(h)) return rc;
    if (qubit < 0 || qubit >= h->n) return fail(DESV_EINVAL, "Qubit " + std::to_string(qubit) + " out of range");
    if (outcome != 0 && outcome != 1) return fail(DESV_EINVAL, "outcome must be 0 or 1");
    const int bit = h->n - 1 - qubit;
    HIP_TRY(h->dtype == DESV_C64 ? launch_collapse<float>(h->sv, h->dim, bit, outcome, scale, h->stream)
                                 : launch_collapse<double>(h->sv, h->dim, bit, outcome, scale, h->stream));
    return DESV_OK;
}

int desv_run_ops(desv_handle h, const double *ops, int64_t n_ops) {
    if (int rc = check_handle(h)) return rc;
    if (n_ops < 0 || (n_ops > 0 && !ops)) return fail(DESV_EINVAL, "bad op table");
    HIP_TRY(hipSetDevice(h->device));
    std::vector<Gate> gates;
    int rc = decode_ops(ops, n_ops, h->n, gates);
    if (rc) return rc;
    return apply_gates(h, gates);
}

int desv_apply_matrix(desv_handle h, const double *mat, const int *qubits, int k) {
    if (int rc = check_handle(h)) return rc;
    if (!mat || !qubits) return fail(DESV_EINVAL, "null matrix or qubit list");
    HIP_TRY(hipSetDevice(h->device));
    Gate g;
    std::string err;
    if (!decode_matrix(mat, qubits, k, h->n, g, err)) return fail(DESV_EINVAL, err);
    std::vector<Gate> gs{g};
    return apply_gates(h, gs);
}

int desv_default_opts(int dtype, desv_compile_opts *opts) {
    if (!opts) return fail(DESV_EINVAL, "null opts");
    int dev = 0;
    (void)hipGetDevice(&dev);
    CompileOpts o = default_opts(dtype, dev, 1);
    opts->fuse_bits = o.fuse_bits;
    opts->tile_bits = o.tile_bits;
    opts->low_bits = o.low_bits;
    opts->per_gate = o.per_gate;
    opts->use_mfma = o.use_mfma;
    opts->dtype = dtype;
    opts->compensated = o.compensated;
    opts->reg_groups = o.reg_groups;
    opts->plan_search = -1;
    return DESV_OK;
}

int desv_program_create(int n, desv_program *out) {
    if (!out) return fail(DESV_EINVAL, "null output pointer");
    if (n < 1 || n > 34) return fail(DESV_EINVAL, "n_qubits must be in [1, 34], got " + std::to_string(n));
    desv_prog *p = new (std::nothrow) desv_prog();
    if (!p) return fail(DESV_ENOMEM, "host allocation failed");
    p->n = n;
    *out = p;
    return DESV_OK;
}

int desv_program_destroy(desv_program p) {
    if (!p) return DESV_OK;
    p->dev.release();
    delete p;
    return DESV_OK;
}

int desv_program_add_ops(desv_program p, const double *ops, int64_t n_ops) {
    if (!p) return fail(DESV_EINVAL, "null program");
    if (n_ops < 0 || (n_ops > 0 && !ops)) return fail(DESV_EINVAL, "bad op table");
    int rc = decode_ops(ops, n_ops, p->n, p->gates);
    if (rc) return rc;
    p->n_logical += n_ops;
    p->compiled = false;
    return DESV_OK;
}

int desv_program_add_matrix(desv_program p, const double *mat, const int *qubits, int k) {
    if (!p) return fail(DESV_EINVAL, "null program");
    if (!mat || !qubits) return fail(DESV_EINVAL, "null matrix or qubit list");
    Gate g;
    std::string err;
    if (!decode_matrix(mat, qubits, k, p->n, g, err)) return fail(DESV_EINVAL, err);
    p->gates.push_back(g);
    p->n_logical += 1;
    p->compiled = false;
    return DESV_OK;
}

int desv_program_compile(desv_program p, const desv_compile_opts *opts) {
    if (!p) return fail(DESV_EINVAL, "null program");
    int dev = 0;
    (void)hipGetDevice(&dev);
    const int dtype = opts ? opts->dtype : DESV_C64;
    CompileOpts o = default_opts(dtype, dev, 1);
    if (opts) {
        if (opts->fuse_bits) o.fuse_bits = opts->fuse_bits;
        if (opts->tile_bits) o.tile_bits = opts->tile_bits;
        if (opts->low_bits) o.low_bits = opts->low_bits;
        o.per_gate = opts->per_gate;
        // MFMA blocks are complex64 only: FP64 MFMA has the same MI355X peak as
        // FP64 VALU (78.6 TF), so complex128 blocks run on VALU
        if (opts->use_mfma && dtype != DESV_C64)
            return fail(DESV_EINVAL, "use_mfma applies to complex64 programs only (complex128 blocks run on VALU)");
        o.use_mfma = opts->use_mfma;
        o.compensated = dtype == DESV_C64 ? opts->compensated : 0;
        if (opts->reg_groups && !(dtype == DESV_C64 && opts->use_mfma))
            return fail(DESV_EINVAL, "reg_groups applies to complex64 programs with use_mfma");
        if (opts->reg_groups < 0 || opts->reg_groups > 2) return fail(DESV_EINVAL, "reg_groups must be 0, 1 or 2");
        o.reg_groups = opts->reg_groups;
    }
    if (o.fuse_bits < 1 || o.fuse_bits > 5) return fail(DESV_EINVAL, "fuse_bits must be in [1, 5]");
    const size_t lim = stage_lds_limit(dev);
    if (((size_t)1 << std::min(o.tile_bits, p->n)) * amp_bytes(dtype) > lim)
        return fail(DESV_EINVAL, "tile_bits exceeds the LDS of this device");
    // wave-tile programs end with an out-of-place relabelling stage instead of
    // restore passes when that stage can run as a wave tile
    o.relabel = o.reg_groups == 2;
    // wave-tile programs (deep circuits) also take the cheapest of up to 64
    // randomised planner passes (about 1.5 ms each for 348 gates on 30
    // qubits; the default budget keeps a long program's compile near 4 of
    // its greedy passes' time beyond 512 gates x 64 passes)
    const int64_t ng = (int64_t)p->gates.size();
    const int auto_search = o.reg_groups == 2 ? (int)std::max<int64_t>(4, std::min<int64_t>(64, 32768 / std::max<int64_t>(ng, 1))) : 0;
    o.search = !opts || opts->plan_search < 0 ? auto_search : opts->plan_search;
    std::string err;
    if (!build_plan(p->gates, p->n, o, p->plan, err)) return fail(DESV_EINVAL, err);
    if (p->plan.fb_from >= 0) {
        RegProgram rp;
        build_wave_program(p->plan, o.compensated != 0, rp);
        if (!rp.relabel_ok) {
            o.relabel = 0;
            if (!build_plan(p->gates, p->n, o, p->plan, err)) return fail(DESV_EINVAL, err);
        }
    }
    p->plan.n_gates = p->n_logical;
    p->opts = o;
    p->compiled = true;
    p->dev.dtype = -1;  // force re-upload
    return DESV_OK;
}

int desv_program_stats(desv_program p, int64_t *gates, int64_t *passes, int64_t *blocks, int64_t *mfma) {
    if (!p) return fail(DESV_EINVAL, "null program");
    if (!p->compiled) return fail(DESV_EINVAL, "program is not compiled");
    if (gates) *gates = p->n_logical;
    if (passes) *passes = (int64_t)p->plan.stages.size();
    if (blocks) *blocks = (int64_t)p->plan.ops.size();
    if (mfma) *mfma = p->plan.mfma_blocks;
    return DESV_OK;
}

int desv_program_plan_sizes(desv_program p, int64_t *n_stages, int64_t *n_ops, int64_t *n_mats) {
    if (!p) return fail(DESV_EINVAL, "null program");
    if (!p->compiled) return fail(DESV_EINVAL, "program is not compiled");
    if (n_stages) *n_stages = (int64_t)p->plan.stages.size();
    if (n_ops) *n_ops = (int64_t)p->plan.ops.size();
    if (n_mats) *n_mats = (int64_t)p->plan.mats.size();
    return DESV_OK;
}

int desv_program_plan_dump(desv_program p, uint64_t *tile_masks, int64_t *stage_ops, int64_t *ops, double *mats) {
    if (!p) return fail(DESV_EINVAL, "null program");
    if (!p->compiled) return fail(DESV_EINVAL, "program is not compiled");
    const Plan &pl = p->plan;
    for (size_t s = 0; s < pl.stages.size(); ++s) {
        if (tile_masks) tile_masks[s] = pl.stages[s].tile_mask;
        if (stage_ops) {
            int64_t *so = stage_ops + 18 * s;
            so[0] = pl.stages[s].op_begin;
            so[1] = pl.stages[s].op_end;
            for (int j = 0; j < 16; ++j) so[2 + j] = j < pl.stages[s].m ? pl.stages[s].sperm[j] : -1;
        }
    }
    if (ops)
        for (size_t i = 0; i < pl.ops.size(); ++i) {
            const DevOp &d = pl.ops[i];
            int64_t *o = ops + 11 * i;
            o[0] = d.kind; o[1] = d.k; o[2] = d.mat; o[3] = d.slot;
            for (int j = 0; j < 5; ++j) o[4 + j] = d.pos[j];
            o[9] = d.cloc;
            o[10] = (int64_t)d.cglob;
        }
    if (mats)
        for (size_t i = 0; i < pl.mats.size(); ++i) {
            mats[2 * i] = pl.mats[i].real();
            mats[2 * i + 1] = pl.mats[i].imag();
        }
    return DESV_OK;
}

int desv_program_reg_sizes(desv_program p, int compensated, int64_t *n_words, int64_t *n_frag_floats) {
    if (!p) return fail(DESV_EINVAL, "null program");
    if (!p->compiled) return fail(DESV_EINVAL, "program is not compiled");
    RegProgram rp;
    if (p->opts.reg_groups == 2) build_wave_program(p->plan, compensated != 0, rp);
    else build_reg_program(p->plan, compensated != 0, rp);
    if (n_words) *n_words = (int64_t)rp.words.size();
    if (n_frag_floats) *n_frag_floats = (int64_t)rp.frags.size();
    return DESV_OK;
}

int desv_program_reg_dump(desv_program p, int compensated, int32_t *stage_word, int32_t *stage_frag,
                          int32_t *words, float *frags) {
    if (!p) return fail(DESV_EINVAL, "null program");
    if (!p->compiled) return fail(DESV_EINVAL, "program is not compiled");
    RegProgram rp;
    if (p->opts.reg_groups == 2) build_wave_program(p->plan, compensated != 0, rp);
    else build_reg_program(p->plan, compensated != 0, rp);
    // entries of plan.stages (a relabelled plan's in-place tail follows them
    // in `words` and `frags`: desv_program_fallback_dump)
    const size_t S = p->plan.stages.size();
    if (stage_word) std::copy(rp.stage_word.begin(), rp.stage_word.begin() + (long)S, stage_word);
    if (stage_frag) std::copy(rp.stage_frag.begin(), rp.stage_frag.begin() + (long)S, stage_frag);
    if (words) std::copy(rp.words.begin(), rp.words.end(), words);
    if (frags) std::copy(rp.frags.begin(), rp.frags.end(), frags);
    return DESV_OK;
}

int desv_program_tail_dump(desv_program p, int compensated, int64_t *fb_from, int64_t *n_fb, uint64_t *store_masks,
                           int32_t *omaps, uint64_t *fb_tile_masks, int64_t *fb_stage_ops, int32_t *fb_stage_word,
                           int32_t *fb_stage_frag) {
    if (!p) return fail(DESV_EINVAL, "null program");
    if (!p->compiled) return fail(DESV_EINVAL, "program is not compiled");
    const Plan &pl = p->plan;
    const size_t S = pl.stages.size(), F = pl.fb_stages.size();
    if (fb_from) *fb_from = pl.fb_from;
    if (n_fb) *n_fb = (int64_t)F;
    for (size_t s = 0; s < S; ++s) {
        const StageRelabel *r = pl.relabel.empty() ? nullptr : &pl.relabel[s];
        if (store_masks) store_masks[s] = r ? r->store_mask : 0;
        if (omaps)
            for (int i = 0; i < 64; ++i)
                omaps[64 * s + i] = r && r->store_mask && i < (int)r->omap.size() ? r->omap[i] : -1;
    }
    for (size_t k = 0; k < F; ++k) {
        const DevStage &st = pl.fb_stages[k];
        if (fb_tile_masks) fb_tile_masks[k] = st.tile_mask;
        if (fb_stage_ops) {
            int64_t *so = fb_stage_ops + 18 * k;
            so[0] = st.op_begin;
            so[1] = st.op_end;
            for (int j = 0; j < 16; ++j) so[2 + j] = j < st.m ? st.sperm[j] : -1;
        }
    }
    if ((fb_stage_word || fb_stage_frag) && F) {
        RegProgram rp;
        if (p->opts.reg_groups == 2) build_wave_program(pl, compensated != 0, rp);
        for (size_t k = 0; k < F; ++k) {
            const bool has = rp.stage_word.size() == S + F;
            if (fb_stage_word) fb_stage_word[k] = has ? rp.stage_word[S + k] : -1;
            if (fb_stage_frag) fb_stage_frag[k] = has ? rp.stage_frag[S + k] : 0;
        }
    }
    return DESV_OK;
}

namespace {
// Compile with defaults if needed and upload the plan for this state.
int prepare_program(desv_handle h, desv_program p) {
    if (int rc = check_handle(h)) return rc;
    if (!p) return fail(DESV_EINVAL, "null program");
    if (p->n != h->n) return fail(DESV_EINVAL, "program and state have different qubit counts");
    HIP_TRY(hipSetDevice(h->device));
    if (!p->compiled) {
        desv_compile_opts o;
        desv_default_opts(h->dtype, &o);
        int rc = desv_program_compile(p, &o);
        if (rc) return rc;
    }
    if (p->dev.dtype != h->dtype || p->dev.device != h->device) {
        int rc = upload_plan(p->plan, h->dtype, h->device, h->stream, p->dev, p->opts.compensated != 0,
                             p->opts.use_mfma ? p->opts.reg_groups : 0);
        if (rc) return rc;
    }
    return DESV_OK;
}
}  // namespace

int desv_run_program(desv_handle h, desv_program p) {
    if (int rc = prepare_program(h, p)) return rc;
    return run_plan_on(h, p->plan, p->dev, 1, h->sv, nullptr, 0);
}

int desv_run_program_timed(desv_handle h, desv_program p, float *stage_ms) {
    if (!stage_ms) return fail(DESV_EINVAL, "null stage_ms");
    if (int rc = prepare_program(h, p)) return rc;
    // the in-place tail of a relabelled plan may run instead (more stages):
    // its extra stages' times are added to the last entry
    const size_t S = p->plan.stages.size();
    const size_t cap = std::max(S, (p->plan.fb_from >= 0 ? (size_t)p->plan.fb_from : S) + p->plan.fb_stages.size());
    std::vector<hipEvent_t> evs(cap + 1, nullptr);
    int rc = DESV_OK;
    for (auto &e : evs)
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            rc = fail(DESV_EDEVICE, "hipEventCreate failed");
            break;
        }
    size_t ran = 0;
    if (!rc) rc = run_plan_on(h, p->plan, p->dev, 1, h->sv, nullptr, 0, evs.data(), &ran);
    if (!rc && hipEventSynchronize(evs[ran]) != hipSuccess) rc = fail(DESV_EDEVICE, "hipEventSynchronize failed");
    for (size_t i = 0; i < S; ++i) stage_ms[i] = 0.f;
    for (size_t i = 0; !rc && i < ran; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, evs[i], evs[i + 1]) != hipSuccess)
            rc = fail(DESV_EDEVICE, "hipEventElapsedTime failed");
        if (S) stage_ms[std::min(i, S - 1)] += ms;
    }
    for (auto &e : evs)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

int desv_pauli_expect(desv_handle h, const uint64_t *xmask, const uint64_t *zmask, int64_t T, double *out) {
    if (int rc = check_handle(h)) return rc;
    if (T == 0) return DESV_OK;
    if (T < 0 || !xmask || !zmask || !out) return fail(DESV_EINVAL, "bad Pauli term arrays");
    if (int rc = check_masks(h, xmask, zmask, T)) return rc;
    HIP_TRY(hipSetDevice(h->device));
    std::vector<int64_t> order(T);
    for (int64_t i = 0; i < T; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return xmask[a] < xmask[b]; });
    std::vector<uint64_t> z(T);
    std::vector<PauliItem> items;
    for (int64_t r = 0; r < T; ++r) {
        const int64_t t = order[r];
        z[r] = zmask[t];
        if (items.empty() || items.back().x != xmask[t] || items.back().count == kTermsPerItem)
            items.push_back({xmask[t], (int32_t)r, 1});
        else
            items.back().count++;
    }
    const size_t ib = items.size() * sizeof(PauliItem), ia = (ib + 255) & ~size_t(255);
    int rc = ensure_small(h, ia + T * sizeof(uint64_t));
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<unsigned char> host(ia + T * sizeof(uint64_t));
    std::memcpy(host.data(), items.data(), ib);
    std::memcpy(host.data() + ia, z.data(), T * sizeof(uint64_t));
    HIP_TRY(hipMemcpyAsync(h->d_small, host.data(), host.size(), hipMemcpyHostToDevice, h->stream));
    int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(h->dim / (kThreads * 8), 256));
    rc = ensure_part(h, 2 * (size_t)T * nblk);
    if (rc) return rc;
    unsigned char *base = reinterpret_cast<unsigned char *>(h->d_small);
    hipError_t e = h->dtype == DESV_C64
                       ? launch_pauli_expect<float>(h->sv, h->n, (const PauliItem *)base, (int)items.size(),
                                                    (const uint64_t *)(base + ia), h->d_part, nblk, h->stream)
                       : launch_pauli_expect<double>(h->sv, h->n, (const PauliItem *)base, (int)items.size(),
                                                     (const uint64_t *)(base + ia), h->d_part, nblk, h->stream);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    std::vector<double> sums(2 * T);
    rc = sum_to_host(h, (int)(2 * T), nblk, sums.data());
    if (rc) return rc;
    for (int64_t r = 0; r < T; ++r) {
        const int64_t t = order[r];
        const int ny = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
        const double re = sums[2 * r], im = sums[2 * r + 1];
        // Re(i^ny (re + i im))
        double v = ny == 0 ? re : ny == 1 ? -im : ny == 2 ? -re : im;
        out[t] = v;
    }
    return DESV_OK;
}

int desv_pauli_matvec(desv_handle out, desv_handle in, const uint64_t *xmask, const uint64_t *zmask,
                      const double *cre, const double *cim, int64_t T) {
    if (int rc = check_handle(out)) return rc;
    if (int rc = check_handle(in)) return rc;
    if (out == in) return fail(DESV_EINVAL, "pauli_matvec needs distinct input and output states");
    if (out->n != in->n || out->dtype != in->dtype) return fail(DESV_EINVAL, "matvec between different shapes");
    if (T < 0 || (T > 0 && (!xmask || !zmask || !cre))) return fail(DESV_EINVAL, "bad Pauli term arrays");
    if (int rc = check_masks(in, xmask, zmask, T)) return rc;
    HIP_TRY(hipSetDevice(out->device));
    HIP_TRY(hipStreamSynchronize(in->stream));
    if (T == 0) {
        HIP_TRY(out->dtype == DESV_C64 ? launch_scale<float>(out->sv, out->dim, 0.0, 0.0, out->stream)
                                       : launch_scale<double>(out->sv, out->dim, 0.0, 0.0, out->stream));
        return DESV_OK;
    }
    return pauli_matvec_raw(out, out->sv, in->sv, xmask, zmask, cre, cim, T);
}

int desv_pauli_sum_expect(desv_handle h, const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                          int64_t T, double *out) {
    if (int rc = check_handle(h)) return rc;
    if (!out) return fail(DESV_EINVAL, "null output");
    if (T < 0 || (T > 0 && (!xmask || !zmask || !coeff))) return fail(DESV_EINVAL, "bad Pauli term arrays");
    if (int rc = check_masks(h, xmask, zmask, T)) return rc;
    HIP_TRY(hipSetDevice(h->device));
    if (T == 0) {
        *out = 0.0;
        return DESV_OK;
    }
    const int64_t grid = (1ll << h->n) >> matvec_tile_bits(h->n);
    if (int rc = ensure_part(h, (size_t)grid)) return rc;
    int64_t g = 0;
    if (int rc = pauli_tiles_raw(h, nullptr, h->sv, xmask, zmask, coeff, nullptr, T, h->d_part, &g)) return rc;
    return sum_to_host(h, 1, (int)g, out);
}

int desv_dense_matvec(desv_handle out, desv_handle in, const double *H) {
    if (int rc = check_handle(out)) return rc;
    if (int rc = check_handle(in)) return rc;
    if (out == in) return fail(DESV_EINVAL, "dense_matvec needs distinct input and output states");
    if (out->n != in->n || out->dtype != in->dtype) return fail(DESV_EINVAL, "matvec between different shapes");
    if (!H) return fail(DESV_EINVAL, "null matrix");
    HIP_TRY(hipSetDevice(out->device));
    HIP_TRY(hipStreamSynchronize(in->stream));
    return dense_matvec_raw(out, out->sv, in->sv, H);
}

// Handles of one shape for a Krylov-basis call; waits for every handle's
// pending work except on `own` (the stream the call launches on).
static int basis_check(desv_handle ref, const desv_handle *hs, int m, hipStream_t own) {
    if (m < 1 || m > kMaxBasis)
        return fail(DESV_EINVAL, "basis calls take 1.." + std::to_string(kMaxBasis) + " vectors, got " +
                                     std::to_string(m));
    if (!hs) return fail(DESV_EINVAL, "null handle array");
    for (int j = 0; j < m; ++j) {
        if (!hs[j]) return fail(DESV_EINVAL, "null handle in basis array");
        if (hs[j]->n != ref->n || hs[j]->dtype != ref->dtype || hs[j]->device != ref->device)
            return fail(DESV_EINVAL, "basis vectors must share n, dtype and device");
        if (hs[j]->stream != own) HIP_TRY(hipStreamSynchronize(hs[j]->stream));
    }
    return DESV_OK;
}

int desv_inner_many(desv_handle w, const desv_handle *vs, int m, double *out) {
    if (int rc = check_handle(w)) return rc;
    if (!out) return fail(DESV_EINVAL, "null output buffer");
    HIP_TRY(hipSetDevice(w->device));
    if (int rc = basis_check(w, vs, m, w->stream)) return rc;
    BasisPtrs ptrs{};
    for (int j = 0; j < m; ++j) ptrs.p[j] = vs[j]->sv;
    const int nblk = reduce_blocks(w->dim);
    if (int rc = ensure_part(w, 2 * (size_t)m * nblk)) return rc;
    hipError_t e = w->dtype == DESV_C64 ? launch_inner_many<float>(w->sv, ptrs, m, w->dim, w->d_part, nblk, w->stream)
                                        : launch_inner_many<double>(w->sv, ptrs, m, w->dim, w->d_part, nblk, w->stream);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    return sum_to_host(w, 2 * m, nblk, out);
}

int desv_combine(const desv_handle *outs, int p, const desv_handle *ins, int m, const double *coef) {
    if (!outs || p < 1 || !outs[0]) return fail(DESV_EINVAL, "combine needs at least one output handle");
    if (!coef) return fail(DESV_EINVAL, "null coefficient array");
    desv_handle h = outs[0];
    HIP_TRY(hipSetDevice(h->device));
    if (int rc = basis_check(h, outs, p, h->stream)) return rc;
    if (int rc = basis_check(h, ins, m, h->stream)) return rc;
    BasisPtrs in{}, out{};
    for (int j = 0; j < m; ++j) in.p[j] = ins[j]->sv;
    for (int l = 0; l < p; ++l) out.p[l] = outs[l]->sv;
    const size_t cb = (size_t)m * p * 2 * sizeof(double);
    if (int rc = ensure_small(h, cb)) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));  // d_small may still be read
    HIP_TRY(hipMemcpyAsync(h->d_small, coef, cb, hipMemcpyHostToDevice, h->stream));
    hipError_t e = h->dtype == DESV_C64
                       ? launch_combine<float>(in, out, m, p, (const double2 *)h->d_small, h->dim, h->stream)
                       : launch_combine<double>(in, out, m, p, (const double2 *)h->d_small, h->dim, h->stream);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    // the outputs other than outs[0] are read next on their own streams
    HIP_TRY(hipStreamSynchronize(h->stream));
    return DESV_OK;
}

int desv_fill_random(desv_handle h, uint64_t seed) {
    if (int rc = check_handle(h)) return rc;
    HIP_TRY(hipSetDevice(h->device));
    hipError_t e = h->dtype == DESV_C64 ? launch_fill_random<float>(h->sv, h->dim, seed, h->stream)
                                        : launch_fill_random<double>(h->sv, h->dim, seed, h->stream);
    if (e != hipSuccess) return fail(DESV_EDEVICE, hipGetErrorString(e));
    return DESV_OK;
}

int desv_adjoint(desv_handle h, const double *ops, int64_t n_ops, const int64_t *param_of_op, int64_t n_params,
                 const uint64_t *xmask, const uint64_t *zmask, const double *coeff, int64_t T,
                 const double *dense_h, double *energy, double *grad) {
    if (int rc = check_handle(h)) return rc;
    if (n_ops < 0 || (n_ops > 0 && (!ops || !param_of_op))) return fail(DESV_EINVAL, "bad op table");
    if (n_params < 0 || (n_params > 0 && !grad)) return fail(DESV_EINVAL, "bad gradient buffer");
    if (!xmask && !dense_h) return fail(DESV_EINVAL, "no Hamiltonian given");
    if (xmask && (T < 0 || (T > 0 && (!zmask || !coeff)))) return fail(DESV_EINVAL, "bad Pauli term arrays");
    if (xmask)
        if (int rc = check_masks(h, xmask, zmask, T)) return rc;
    HIP_TRY(hipSetDevice(h->device));
    // forward gates + reverse list with gradient probes
    std::vector<Gate> fwd;
    std::vector<Gate> rev;
    std::vector<int64_t> slot_param;
    {
        std::string err;
        std::vector<std::vector<Gate>> per_row(n_ops);
        for (int64_t i = 0; i < n_ops; ++i) {
            if (!decode_row(ops + 4 * i, h->n, per_row[i], err)) return fail(DESV_EINVAL, err);
            for (auto &g : per_row[i]) fwd.push_back(g);
        }
        for (int64_t i = n_ops - 1; i >= 0; --i) {
            const int64_t pi = param_of_op[i];
            if (pi >= n_params) return fail(DESV_EINVAL, "param_of_op index out of range");
            if (pi >= 0) {
                Gate gg;
                if (!grad_gate_for_row(ops + 4 * i, h->n, (int)slot_param.size(), gg, err))
                    return fail(DESV_EINVAL, err);
                slot_param.push_back(pi);
                rev.push_back(gg);
            }
            for (auto it = per_row[i].rbegin(); it != per_row[i].rend(); ++it) rev.push_back(inverse_gate(*it));
        }
    }
    int rc = ensure_aux(h, 0);
    if (rc) return rc;
    rc = ensure_aux(h, 1);
    if (rc) return rc;
    // forward
    rc = apply_gates(h, fwd);
    if (rc) return rc;
    // lambda = H psi
    if (xmask) {
        std::vector<double> zero(T, 0.0);
        rc = pauli_matvec_raw(h, h->aux[1], h->sv, xmask, zmask, coeff, zero.data(), T);
    } else {
        rc = dense_matvec_raw(h, h->aux[1], h->sv, dense_h);
    }
    if (rc) return rc;
    double e_re = 0;
    rc = inner_raw(h, h->sv, h->aux[1], &e_re, nullptr);
    if (rc) return rc;
    if (energy) *energy = e_re;
    for (int64_t k = 0; k < n_params; ++k) grad[k] = 0.0;
    if (slot_param.empty()) return DESV_OK;
    // reverse sweep on a copy of psi (the handle keeps the final state)
    HIP_TRY(h->bytes % 16 == 0 ? launch_copy16(h->aux[0], h->sv, h->bytes, h->stream)
                               : hipMemcpyAsync(h->aux[0], h->sv, h->bytes, hipMemcpyDeviceToDevice, h->stream));
    CompileOpts o = default_opts(h->dtype, h->device, 2);
    Plan plan;
    std::string err;
    if (!build_plan(rev, h->n, o, plan, err)) return fail(DESV_EINVAL, err);
    const int64_t pstride = 1ll << (h->n - std::min(h->n, o.tile_bits));
    rc = ensure_part(h, (size_t)plan.n_grad_slots * pstride);
    if (rc) return rc;
    rc = upload_plan(plan, h->dtype, h->device, h->stream, h->scratch_plan, h->compensated != 0);
    if (rc) return rc;
    rc = run_plan_on(h, plan, h->scratch_plan, 2, h->aux[0], h->aux[1], pstride);
    if (rc) return rc;
    std::vector<double> slots(plan.n_grad_slots);
    rc = sum_to_host(h, plan.n_grad_slots, (int)pstride, slots.data());
    if (rc) return rc;
    for (size_t s = 0; s < slot_param.size(); ++s) grad[slot_param[s]] += slots[s];
    return DESV_OK;
}

}  // extern "C"
