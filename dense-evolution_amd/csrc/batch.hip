// Heterogeneous VQE batches (desv_batch_*): many (circuit, Hamiltonian,
// theta) instances of uneven qubit counts evaluated together.
//
// Every instance is compiled once (forward plan, reverse plan with gradient
// probes); new angles only refresh the fused matrices on the host.  All device
// data lives in one arena, and one kernel launch covers stage k of every
// instance that has a stage k (tasks looked up per workgroup), so a whole
// batch is ~(forward stages + reverse stages + 3) launches, captured once into
// a hipGraph and replayed.  Generalises run_batch_jit
// (dense_evolution/backends/statevector.py:496-584), which vmaps one circuit
// over parameter rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/desv.h"
#include "desv_internal.h"
#include "kernels.h"

using namespace desv;

namespace {

inline size_t abytes(int dtype) { return dtype == DESV_C64 ? 8 : 16; }
inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct BInst {
    int n = 0;
    std::vector<double> rows;
    std::vector<int64_t> pmap;
    int64_t n_params = 0;
    std::vector<uint64_t> x, z;
    std::vector<double> coeff;
    bool dirty = true;
    std::vector<unsigned char> init;  // optional initial state (host copy)
    // compiled
    std::vector<Gate> fg, rg;
    std::vector<int> rg_src;          // >= 0: inverse of fg[i]; < 0: probe
    std::vector<int64_t> slot_param;
    Plan fp, rp;
    int mf = 0, mr = 0;
    // device (arena offsets resolved to pointers)
    void *psi = nullptr, *psi2 = nullptr, *lam = nullptr, *init_dev = nullptr;
    DevOp *fops = nullptr, *rops = nullptr;
    size_t fmat = 0, rmat = 0;        // element offsets into the matrix arena
    double *gpart = nullptr, *epart = nullptr;
    int64_t pstride = 0, mv_wg = 0, mv_chunk = 0;
    void *groups = nullptr;
    uint64_t *zz = nullptr;
    double2 *cc = nullptr;
    int ngroups = 0;
    int e_out = 0, s_out = 0;
};

struct Launch {
    int kind;        // 0 fill, 1 stage, 2 matvec, 3 sum, 4 start mark
    int nv = 1;
    int heavy = 0;  // stage kernel variant
    const void *tasks = nullptr;
    const int32_t *wg_task = nullptr;  // stage launches: task of each workgroup
    int ntasks = 0;
    int64_t grid = 0;
    int max_m = 0;
};

int bfail(int code, const std::string &m) { return set_last_error(code, m); }

#define BHIP(expr)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return bfail(e_ == hipErrorOutOfMemory ? DESV_ENOMEM : DESV_EDEVICE,                 \
                         std::string(#expr) + ": " + hipGetErrorString(e_));                    \
    } while (0)

template <typename F> void parallel_for(size_t n, F f) {
    unsigned hw = std::thread::hardware_concurrency();
    size_t nt = std::min<size_t>(std::max(1u, std::min(hw, 16u)), n / 16 + 1);
    if (nt <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (size_t i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

}  // namespace

struct desv_batch_s {
    int dtype = 0;
    int device = 0;
    int tile_bits = 0;
    int fuse_bits = 2;
    int use_mfma = 0;
    int use_graph = 1;
    // complex64: matrices as FP32 hi + lo pairs (pool doubled: lo part
    // mat_entries elements after the hi part), applied lo part first
    int compensated = 0;
    size_t mat_entries = 0;
    // forward only: no Hamiltonian, no reverse sweep; a run applies each
    // instance's circuit to |0...0> (or its initial state) and the final
    // states sit back to back in the arena (desv_batch_get_states)
    int forward_only = 0;
    void *states_dev = nullptr;
    size_t states_bytes = 0;
    double last_us = 0;
    std::vector<BInst> inst;
    bool built = false;
    hipStream_t stream = nullptr;
    void *arena = nullptr;
    size_t arena_bytes = 0;
    void *mats_dev = nullptr;
    void *mats_host = nullptr;  // pinned
    size_t mats_bytes = 0;
    double *out_dev = nullptr, *out_host = nullptr;  // pinned host
    int n_out = 0;
    // device wall-clock stamps of the last run: [0] start, [1 + i] the time
    // instance i's last result was summed
    unsigned long long *stamps_dev = nullptr;
    int64_t n_stamps = 0;
    // launch lists: energies + gradients, energies only (forward, lambda,
    // energy sums), and the reverse sweep alone (gradients of the state an
    // energies-only run left: psi copy and lambda), each with its graph
    std::vector<Launch> full, energy_only, reverse_only;
    hipGraphExec_t graph_full = nullptr, graph_energy = nullptr, graph_reverse = nullptr;
    // the device holds the psi copy and lambda of an energies-only run made
    // with the current matrices and initial states
    bool reverse_ready = false;
    std::string err;

    void release() {
        if (graph_full) (void)hipGraphExecDestroy(graph_full);
        if (graph_energy) (void)hipGraphExecDestroy(graph_energy);
        if (graph_reverse) (void)hipGraphExecDestroy(graph_reverse);
        graph_full = graph_energy = graph_reverse = nullptr;
        reverse_ready = false;
        if (arena) (void)hipFree(arena);
        if (mats_host) (void)hipHostFree(mats_host);
        if (out_host) (void)hipHostFree(out_host);
        arena = mats_host = nullptr;
        out_host = nullptr;
        full.clear();
        energy_only.clear();
        reverse_only.clear();
        built = false;
    }
};

namespace {

int choose_m(desv_batch b, int n, int nv) {
    const size_t lim = stage_lds_limit(b->device);
    int m = b->tile_bits > 0 ? b->tile_bits : (b->dtype == DESV_C64 ? 11 : 10);
    // a lone instance fills the GPU only with enough tiles: keep >= 2^10
    // workgroups (4 per CU) unless the tile would drop below 2^8
    if (b->tile_bits == 0 && b->inst.size() == 1) m = std::min(m, std::max(8, n - 10));
    while (m > 6 && (size_t)nv * ((size_t)1 << m) * abytes(b->dtype) > lim) --m;
    return std::min(n, m);
}

// forward gates from the (patched) rows; reverse list = per row, from last
// to first: [probe if theta-carrying] + inverses of the row's gates.
int decode_instance(BInst &in, bool structure) {
    std::string err;
    const int64_t nops = (int64_t)in.pmap.size();
    std::vector<std::vector<Gate>> per(nops);
    std::vector<Gate> fg;
    for (int64_t i = 0; i < nops; ++i) {
        if (!decode_row(in.rows.data() + 4 * i, in.n, per[i], err)) return bfail(DESV_EINVAL, err);
        for (auto &g : per[i]) fg.push_back(g);
    }
    std::vector<Gate> rg;
    std::vector<int64_t> slot_param;
    int slot = 0;
    for (int64_t i = nops - 1; i >= 0; --i) {
        if (in.pmap[i] >= 0) {
            Gate gg;
            if (!grad_gate_for_row(in.rows.data() + 4 * i, in.n, slot, gg, err)) return bfail(DESV_EINVAL, err);
            slot_param.push_back(in.pmap[i]);
            ++slot;
            rg.push_back(gg);
        }
        for (auto it = per[i].rbegin(); it != per[i].rend(); ++it) rg.push_back(inverse_gate(*it));
    }
    in.fg.swap(fg);
    in.rg.swap(rg);
    if (structure) in.slot_param.swap(slot_param);
    return DESV_OK;
}

// lo_entries > 0: compensated pool, the lo parts lo_entries elements after
// the hi parts
void write_mats(const Plan &p, int dtype, void *host, size_t off, size_t lo_entries) {
    if (dtype == DESV_C64) {
        float *hi = reinterpret_cast<float *>(host) + 2 * off;
        round_mats_f32(p, hi);
        if (lo_entries) round_mats_f32_lo(p, hi, hi + 2 * lo_entries);
    } else {
        double *f = reinterpret_cast<double *>(host) + 2 * off;
        for (size_t i = 0; i < p.mats.size(); ++i) {
            f[2 * i] = p.mats[i].real();
            f[2 * i + 1] = p.mats[i].imag();
        }
    }
}

int run_launches(desv_batch b, const std::vector<Launch> &ls) {
    for (const Launch &l : ls) {
        hipError_t e = hipSuccess;
        const bool f = b->dtype == DESV_C64;
        StageTask dummy;
        std::memset(&dummy, 0, sizeof(dummy));
        switch (l.kind) {
            case 0:
                e = f ? launch_fill_tasks<float>((const FillTask *)l.tasks, l.ntasks, l.grid, b->stream)
                      : launch_fill_tasks<double>((const FillTask *)l.tasks, l.ntasks, l.grid, b->stream);
                break;
            case 1:
                e = f ? launch_stage<float>(l.nv, l.heavy, b->mat_entries != 0, (const StageTask *)l.tasks, l.wg_task,
                                            dummy, l.grid, l.max_m, b->stream)
                      : launch_stage<double>(l.nv, l.heavy, false, (const StageTask *)l.tasks, l.wg_task, dummy, l.grid,
                                             l.max_m, b->stream);
                break;
            case 2:
                e = f ? launch_matvec_tasks<float>((const MatvecTask *)l.tasks, l.ntasks, l.grid, b->stream)
                      : launch_matvec_tasks<double>((const MatvecTask *)l.tasks, l.ntasks, l.grid, b->stream);
                break;
            case 3: e = launch_sum_tasks((const SumTask *)l.tasks, l.ntasks, b->stream); break;
            case 4: e = launch_mark(b->stamps_dev, b->n_stamps, b->stream); break;
        }
        if (e != hipSuccess) return bfail(DESV_EDEVICE, std::string("batch launch: ") + hipGetErrorString(e));
    }
    return DESV_OK;
}

// Byte signature of an instance's gate structure (qubit count and, per gate of
// the forward and reverse lists, kind, width, bits, controls and probe
// slot): instances with equal signatures differ only in their matrices, so
// they share one compiled plan whose matrices are recomputed from their own
// gates.
std::string gate_signature(const BInst &in) {
    std::string s;
    auto put = [&](const void *p, size_t nb) { s.append(reinterpret_cast<const char *>(p), nb); };
    put(&in.n, sizeof(in.n));
    for (const auto *list : {&in.fg, &in.rg}) {
        const size_t cnt = list->size();
        put(&cnt, sizeof(cnt));
        for (const Gate &g : *list) {
            const int hdr[5] = {g.kind, g.k, g.grad_diag ? 1 : 0, g.grad_slot, (int)g.m.size()};
            put(hdr, sizeof(hdr));
            put(g.bits, sizeof(g.bits));
            put(&g.cmask, sizeof(g.cmask));
        }
    }
    return s;
}

int build(desv_batch b) {
    b->release();
    BHIP(hipSetDevice(b->device));
    const size_t ab = abytes(b->dtype);
    // host compile: decode every instance, compile one plan per distinct gate
    // structure, then give the other instances a copy with their own matrices
    std::vector<int> rc(b->inst.size(), 0);
    std::vector<std::string> errs(b->inst.size());
    parallel_for(b->inst.size(), [&](size_t i) {
        if (decode_instance(b->inst[i], true)) {
            rc[i] = DESV_EINVAL;
            errs[i] = desv_last_error();
        }
    });
    for (size_t i = 0; i < rc.size(); ++i)
        if (rc[i]) return bfail(rc[i], "instance " + std::to_string(i) + ": " + errs[i]);
    std::vector<int64_t> rep(b->inst.size());
    {
        std::map<std::string, int64_t> first;
        for (size_t i = 0; i < b->inst.size(); ++i)
            rep[i] = first.emplace(gate_signature(b->inst[i]), (int64_t)i).first->second;
    }
    std::vector<size_t> reps;
    for (size_t i = 0; i < b->inst.size(); ++i)
        if (rep[i] == (int64_t)i) reps.push_back(i);
    parallel_for(reps.size(), [&](size_t r) {
        const size_t i = reps[r];
        BInst &in = b->inst[i];
        CompileOpts of, orv;
        of.tile_bits = in.mf = choose_m(b, in.n, 1);
        orv.tile_bits = in.mr = choose_m(b, in.n, 2);
        of.low_bits = orv.low_bits = std::min(5, in.n);
        of.fuse_bits = orv.fuse_bits = b->fuse_bits;
        // MFMA blocks in the forward pass only (complex64); the reverse sweep's
        // gradient probes keep its blocks small
        of.use_mfma = b->dtype == DESV_C64 ? b->use_mfma : 0;
        orv.nv = 2;
        std::string err;
        in.rp = Plan();
        if (!build_plan(in.fg, in.n, of, in.fp, err) ||
            (!b->forward_only && !build_plan(in.rg, in.n, orv, in.rp, err))) {
            rc[i] = DESV_EINVAL;
            errs[i] = err;
            return;
        }
        in.pstride = 1ll << (in.n - in.mr);
        in.dirty = false;
    });
    for (size_t i = 0; i < rc.size(); ++i)
        if (rc[i]) return bfail(rc[i], "instance " + std::to_string(i) + ": " + errs[i]);
    parallel_for(b->inst.size(), [&](size_t i) {
        if (rep[i] == (int64_t)i) return;
        BInst &in = b->inst[i];
        const BInst &r = b->inst[rep[i]];
        in.fp = r.fp;
        in.rp = r.rp;
        in.mf = r.mf;
        in.mr = r.mr;
        in.pstride = r.pstride;
        refresh_plan_mats(in.fp, in.fg);
        if (!b->forward_only) refresh_plan_mats(in.rp, in.rg);
        in.dirty = false;
    });

    // arena layout
    size_t off = 0, moff = 0;
    int n_out = 0;
    struct Offs {
        size_t psi, psi2, lam, init, fops, rops, gpart, epart, groups, zz, cc;
    };
    std::vector<Offs> o(b->inst.size());
    std::vector<std::vector<MatvecGroup>> groups(b->inst.size());
    std::vector<std::vector<uint64_t>> zs(b->inst.size());
    std::vector<std::vector<double2>> ccs(b->inst.size());
    // the final states first, back to back in instance order
    for (size_t i = 0; i < b->inst.size(); ++i) {
        o[i].psi = off;
        off += ((size_t)1 << b->inst[i].n) * ab;
    }
    const size_t states_bytes = off;
    off = align256(off);
    const bool rev = !b->forward_only;
    for (size_t i = 0; i < b->inst.size(); ++i) {
        BInst &in = b->inst[i];
        const size_t sb = ((size_t)1 << in.n) * ab;
        o[i].psi2 = off; off = align256(off + (rev ? sb : 0));
        o[i].lam = off; off = align256(off + (rev ? sb : 0));
        o[i].init = off; off = align256(off + (in.init.empty() ? 0 : sb));
        o[i].fops = off; off = align256(off + in.fp.ops.size() * sizeof(DevOp));
        o[i].rops = off; off = align256(off + in.rp.ops.size() * sizeof(DevOp));
        o[i].gpart = off; off = align256(off + (size_t)std::max(1, in.rp.n_grad_slots) * in.pstride * 8);
        // lambda = H psi: one aligned LDS tile of amplitudes per workgroup
        in.mv_chunk = 1ll << matvec_tile_bits(in.n);
        in.mv_wg = ((1ll << in.n) + in.mv_chunk - 1) / in.mv_chunk;
        o[i].epart = off; off = align256(off + in.mv_wg * 8);
        // Pauli groups (sorted by flip mask), i^#Y folded into the coefficient
        const int64_t T = (int64_t)in.x.size();
        std::vector<int64_t> ord(T);
        for (int64_t t = 0; t < T; ++t) ord[t] = t;
        std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t c) { return in.x[a] < in.x[c]; });
        for (int64_t r = 0; r < T; ++r) {
            const int64_t t = ord[r];
            zs[i].push_back(in.z[t]);
            double a = in.coeff[t], c = 0;
            for (int k = 0; k < (__builtin_popcountll(in.x[t] & in.z[t]) & 3); ++k) {
                const double tmp = a;
                a = -c;
                c = tmp;
            }
            ccs[i].push_back(make_double2(a, c));
            if (groups[i].empty() || groups[i].back().x != in.x[t])
                groups[i].push_back({in.x[t], (int32_t)r, (int32_t)r + 1});
            else
                groups[i].back().end = (int32_t)r + 1;
        }
        in.ngroups = (int)groups[i].size();
        o[i].groups = off; off = align256(off + groups[i].size() * sizeof(MatvecGroup) + 8);
        o[i].zz = off; off = align256(off + zs[i].size() * 8 + 8);
        o[i].cc = off; off = align256(off + ccs[i].size() * sizeof(double2) + 16);
        in.fmat = moff; moff += in.fp.mats.size();
        in.rmat = moff; moff += in.rp.mats.size();
        if (!rev) continue;
        in.e_out = n_out++;
        in.s_out = n_out;
        n_out += in.rp.n_grad_slots;
    }
    const size_t out_off = off;
    off = align256(off + (size_t)std::max(1, n_out) * 8);
    const size_t stamps_off = off;
    off = align256(off + (b->inst.size() + 1) * sizeof(unsigned long long));
    const size_t mats_off = off;
    const bool comp = b->compensated && b->dtype == DESV_C64;
    if (comp && moff > (size_t)0x3fffffff) return bfail(DESV_EINVAL, "matrix pool too large");
    b->mat_entries = comp ? moff : 0;
    b->mats_bytes = std::max<size_t>(moff, 1) * ab * (comp ? 2 : 1);
    off = align256(off + b->mats_bytes);
    // task arrays are appended below; compute them after pointers are known
    const size_t tasks_off = off;

    // launch plans (host task arrays), sizes first
    int maxF = 0, maxR = 0;
    for (auto &in : b->inst) {
        maxF = std::max(maxF, (int)in.fp.stages.size());
        maxR = std::max(maxR, (int)in.rp.stages.size());
    }
    std::vector<unsigned char> task_blob;
    struct PendingLaunch {
        Launch l;
        size_t blob_off;
        size_t wg_off = 0;
    };
    std::vector<PendingLaunch> pend_full, pend_energy, pend_reverse;
    auto push_blob = [&](const void *p, size_t bytes) {
        const size_t at = align256(task_blob.size());
        task_blob.resize(at + bytes);
        std::memcpy(task_blob.data() + at, p, bytes);
        return at;
    };
    char *A = nullptr;  // arena base (device), resolved after allocation
    // We need device pointers inside the tasks, so allocate the arena first.
    size_t est_tasks = 0;
    {
        size_t nt = 0;
        size_t tiles = 0;
        for (auto &in : b->inst) {
            nt += in.fp.stages.size() + in.rp.stages.size() + 3 + in.rp.n_grad_slots;
            for (auto &st : in.fp.stages) tiles += st.n_tiles;
            for (auto &st : in.rp.stages) tiles += st.n_tiles;
        }
        est_tasks = nt * sizeof(StageTask) + tiles * sizeof(int32_t) + (size_t)(maxF + maxR + 8) * 16 * 256 + 4096;
    }
    b->arena_bytes = tasks_off + align256(est_tasks);
    BHIP(hipMalloc(&b->arena, b->arena_bytes));
    A = reinterpret_cast<char *>(b->arena);
    BHIP(hipHostMalloc(&b->mats_host, b->mats_bytes, hipHostMallocDefault));
    BHIP(hipHostMalloc((void **)&b->out_host, (size_t)std::max(1, n_out) * 8, hipHostMallocDefault));
    b->out_dev = reinterpret_cast<double *>(A + out_off);
    b->stamps_dev = reinterpret_cast<unsigned long long *>(A + stamps_off);
    b->states_dev = A;
    b->states_bytes = states_bytes;
    b->n_stamps = (int64_t)b->inst.size();
    b->mats_dev = A + mats_off;
    b->n_out = n_out;
    for (size_t i = 0; i < b->inst.size(); ++i) {
        BInst &in = b->inst[i];
        in.psi = A + o[i].psi;
        in.psi2 = A + o[i].psi2;
        in.lam = A + o[i].lam;
        in.init_dev = in.init.empty() ? nullptr : A + o[i].init;
        in.fops = reinterpret_cast<DevOp *>(A + o[i].fops);
        in.rops = reinterpret_cast<DevOp *>(A + o[i].rops);
        in.gpart = reinterpret_cast<double *>(A + o[i].gpart);
        in.epart = reinterpret_cast<double *>(A + o[i].epart);
        in.groups = A + o[i].groups;
        in.zz = reinterpret_cast<uint64_t *>(A + o[i].zz);
        in.cc = reinterpret_cast<double2 *>(A + o[i].cc);
    }
    // start mark (clears the completion stamps)
    {
        Launch l;
        l.kind = 4;
        pend_full.push_back({l, 0});
        pend_energy.push_back({l, 0});
    }
    // fill
    {
        std::vector<FillTask> ft;
        int64_t wg = 0;
        for (auto &in : b->inst) {
            FillTask t;
            t.dst = in.psi;
            t.src = in.init_dev;
            t.dim = 1ll << in.n;
            t.chunk = std::min<int64_t>(t.dim, 8192);
            t.wg_begin = wg;
            wg += (t.dim + t.chunk - 1) / t.chunk;
            ft.push_back(t);
        }
        Launch l;
        l.kind = 0;
        l.ntasks = (int)ft.size();
        l.grid = wg;
        const size_t at = push_blob(ft.data(), ft.size() * sizeof(FillTask));
        pend_full.push_back({l, at});
        pend_energy.push_back({l, at});
    }
    // the launches of stage k of every instance that has one (one per kernel variant)
    auto stage_launches = [&](bool rev, int k, std::vector<PendingLaunch> &out) {
        for (int heavy = 0; heavy < 4; ++heavy) {
            std::vector<StageTask> ts;
            int64_t wg = 0;
            int max_m = 0;
            for (auto &in : b->inst) {
                const Plan &p = rev ? in.rp : in.fp;
                if ((int)p.stages.size() <= k || p.stages[k].heavy != heavy) continue;
                StageTask t;
                std::memset(&t, 0, sizeof(t));
                t.st = p.stages[k];
                t.n = in.n;
                t.sv0 = rev ? in.psi2 : in.psi;
                t.sv1 = rev ? in.lam : nullptr;
                t.ops = rev ? in.rops : in.fops;
                t.mats = reinterpret_cast<char *>(b->mats_dev) + (rev ? in.rmat : in.fmat) * ab;
                t.lo_off = (int32_t)b->mat_entries;
                t.partials = in.gpart;
                t.pstride = in.pstride;
                t.wg_begin = wg;
                wg += t.st.n_tiles;
                max_m = std::max(max_m, (int)t.st.m);
                ts.push_back(t);
            }
            if (ts.empty()) continue;
            Launch l;
            l.kind = 1;
            l.nv = rev ? 2 : 1;
            l.heavy = heavy;
            l.ntasks = (int)ts.size();
            l.grid = wg;
            l.max_m = max_m;
            std::vector<int32_t> wt((size_t)wg);
            for (size_t ti = 0; ti < ts.size(); ++ti)
                for (int64_t w = 0; w < ts[ti].st.n_tiles; ++w) wt[ts[ti].wg_begin + w] = (int32_t)ti;
            PendingLaunch pl{l, push_blob(ts.data(), ts.size() * sizeof(StageTask))};
            pl.wg_off = push_blob(wt.data(), wt.size() * sizeof(int32_t));
            out.push_back(pl);
        }
    };
    {
        const size_t before = pend_full.size();
        for (int k = 0; k < maxF; ++k) stage_launches(false, k, pend_full);
        for (size_t i = before; i < pend_full.size(); ++i) pend_energy.push_back(pend_full[i]);
    }
    if (rev) {
        std::vector<MatvecTask> mt;
        int64_t wg = 0;
        for (auto &in : b->inst) {
            MatvecTask t;
            t.n = in.n;
            t.ngroups = in.ngroups;
            t.psi = in.psi;
            t.lam = in.lam;
            t.psi2 = in.psi2;
            t.groups = in.groups;
            t.z = in.zz;
            t.cc = in.cc;
            t.epart = in.epart;
            t.wg_begin = wg;
            t.chunk = in.mv_chunk;
            wg += in.mv_wg;
            mt.push_back(t);
        }
        Launch l;
        l.kind = 2;
        l.ntasks = (int)mt.size();
        l.grid = wg;
        const size_t at = push_blob(mt.data(), mt.size() * sizeof(MatvecTask));
        pend_full.push_back({l, at});
        pend_energy.push_back({l, at});
    }
    // Sums of an instance's energy and gradient partials run right after its
    // last reverse stage (or after the matvec, without one), so each instance's
    // results are final, and time-stamped, as soon as its own work is done.
    auto sum_launch = [&](const std::vector<SumTask> &ts, std::vector<PendingLaunch> &out) {
        if (ts.empty()) return;
        Launch l;
        l.kind = 3;
        l.ntasks = (int)ts.size();
        out.push_back({l, push_blob(ts.data(), ts.size() * sizeof(SumTask))});
    };
    auto sums_of = [&](size_t i, bool grads, std::vector<SumTask> &ts) {
        const BInst &in = b->inst[i];
        SumTask t;
        t.src = in.epart;
        t.count = in.mv_wg;
        t.dst = b->out_dev + in.e_out;
        t.stamp = b->stamps_dev + 1 + i;
        ts.push_back(t);
        for (int s = 0; grads && s < in.rp.n_grad_slots; ++s) {
            SumTask g;
            g.src = in.gpart + (int64_t)s * in.pstride;
            g.count = in.pstride;
            g.dst = b->out_dev + in.s_out + s;
            g.stamp = b->stamps_dev + 1 + i;
            ts.push_back(g);
        }
    };
    {
        // the reverse-only list starts with its own mark (completion times of
        // a reverse run count from its first kernel)
        Launch l;
        l.kind = 4;
        pend_reverse.push_back({l, 0});
    }
    for (int k = -1; rev && k < maxR; ++k) {
        const size_t before = pend_full.size();
        if (k >= 0) stage_launches(true, k, pend_full);
        std::vector<SumTask> done;
        for (size_t i = 0; i < b->inst.size(); ++i)
            if ((int)b->inst[i].rp.stages.size() == k + 1) sums_of(i, true, done);
        sum_launch(done, pend_full);
        for (size_t i = before; i < pend_full.size(); ++i) pend_reverse.push_back(pend_full[i]);
    }
    if (rev) {
        std::vector<SumTask> se;
        for (size_t i = 0; i < b->inst.size(); ++i) sums_of(i, false, se);
        sum_launch(se, pend_energy);
    }
    if (task_blob.size() > b->arena_bytes - tasks_off) return bfail(DESV_EDEVICE, "internal: task area too small");
    // upload the static per-instance tables
    auto up = [&](size_t at, const void *src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpy(A + at, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    for (size_t i = 0; i < b->inst.size(); ++i) {
        BInst &in = b->inst[i];
        BHIP(up(o[i].fops, in.fp.ops.data(), in.fp.ops.size() * sizeof(DevOp)));
        BHIP(up(o[i].rops, in.rp.ops.data(), in.rp.ops.size() * sizeof(DevOp)));
        BHIP(up(o[i].init, in.init.data(), in.init.size()));
        BHIP(up(o[i].groups, groups[i].data(), groups[i].size() * sizeof(MatvecGroup)));
        BHIP(up(o[i].zz, zs[i].data(), zs[i].size() * 8));
        BHIP(up(o[i].cc, ccs[i].data(), ccs[i].size() * sizeof(double2)));
    }
    BHIP(hipMemcpy(A + tasks_off, task_blob.data(), task_blob.size(), hipMemcpyHostToDevice));
    for (auto &p : pend_full) {
        p.l.tasks = A + tasks_off + p.blob_off;
        if (p.l.kind == 1) p.l.wg_task = reinterpret_cast<const int32_t *>(A + tasks_off + p.wg_off);
        b->full.push_back(p.l);
    }
    for (auto &p : pend_energy) {
        p.l.tasks = A + tasks_off + p.blob_off;
        if (p.l.kind == 1) p.l.wg_task = reinterpret_cast<const int32_t *>(A + tasks_off + p.wg_off);
        b->energy_only.push_back(p.l);
    }
    for (auto &p : pend_reverse) {
        p.l.tasks = A + tasks_off + p.blob_off;
        if (p.l.kind == 1) p.l.wg_task = reinterpret_cast<const int32_t *>(A + tasks_off + p.wg_off);
        b->reverse_only.push_back(p.l);
    }
    // matrices
    for (auto &in : b->inst) {
        write_mats(in.fp, b->dtype, b->mats_host, in.fmat, b->mat_entries);
        write_mats(in.rp, b->dtype, b->mats_host, in.rmat, b->mat_entries);
    }
    if (!b->stream) BHIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    if (b->use_graph) {
        // the batch counts as built only once both graphs exist; a failed
        // capture or instantiation releases what was made and leaves it unbuilt
        for (int which = 0; which < 3; ++which) {
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) {
                b->release();
                return bfail(DESV_EDEVICE, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
            }
            const int r = run_launches(b, which == 0 ? b->full : which == 1 ? b->energy_only : b->reverse_only);
            e = hipStreamEndCapture(b->stream, &g);
            hipGraphExec_t ex = nullptr;
            if (!r && e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (r || e != hipSuccess) {
                const std::string msg = r ? std::string(desv_last_error())
                                          : std::string("graph capture: ") + hipGetErrorString(e);
                b->release();
                return bfail(r ? r : DESV_EDEVICE, msg);
            }
            (which == 0 ? b->graph_full : which == 1 ? b->graph_energy : b->graph_reverse) = ex;
        }
    }
    b->built = true;
    return DESV_OK;
}

int refresh(desv_batch b) {
    std::vector<int> rc(b->inst.size(), 0);
    parallel_for(b->inst.size(), [&](size_t i) {
        BInst &in = b->inst[i];
        if (!in.dirty) return;
        if (decode_instance(in, false)) {
            rc[i] = 1;
            return;
        }
        refresh_plan_mats(in.fp, in.fg);
        refresh_plan_mats(in.rp, in.rg);
        write_mats(in.fp, b->dtype, b->mats_host, in.fmat, b->mat_entries);
        write_mats(in.rp, b->dtype, b->mats_host, in.rmat, b->mat_entries);
        in.dirty = false;
    });
    for (int r : rc)
        if (r) return bfail(DESV_EINVAL, "refresh failed");
    return DESV_OK;
}

// mode 0: energies only, 1: energies and gradients, 2: the reverse sweep of
// the last energies-only run (gradients; energies re-summed)
int execute(desv_batch b, int mode) {
    if (mode == 2) {
        bool dirty = false;
        for (const auto &in : b->inst) dirty |= in.dirty;
        if (!b->built || !b->reverse_ready || dirty)
            return bfail(DESV_EINVAL, "the reverse sweep needs an energies-only run with the current parameters");
    } else if (!b->built) {
        int r = build(b);
        if (r) return r;
    } else {
        int r = refresh(b);
        if (r) return r;
    }
    BHIP(hipSetDevice(b->device));
    if (mode != 2)
        BHIP(hipMemcpyAsync(b->mats_dev, b->mats_host, b->mats_bytes, hipMemcpyHostToDevice, b->stream));
    b->reverse_ready = false;
    if (b->use_graph) {
        BHIP(hipGraphLaunch(mode == 1 ? b->graph_full : mode == 0 ? b->graph_energy : b->graph_reverse, b->stream));
    } else {
        int r = run_launches(b, mode == 1 ? b->full : mode == 0 ? b->energy_only : b->reverse_only);
        if (r) return r;
    }
    b->reverse_ready = mode == 0 && !b->forward_only;
    BHIP(hipMemcpyAsync(b->out_host, b->out_dev, (size_t)std::max(1, b->n_out) * 8, hipMemcpyDeviceToHost,
                        b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return DESV_OK;
}

}  // namespace

extern "C" {

int desv_batch_create(int dtype, int device, desv_batch *out) {
    if (!out) return bfail(DESV_EINVAL, "null output pointer");
    if (dtype != DESV_C64 && dtype != DESV_C128) return bfail(DESV_EINVAL, "dtype must be DESV_C64 or DESV_C128");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return bfail(DESV_EDEVICE, "no HIP device available");
    }
    if (device < 0 || device >= count) return bfail(DESV_EINVAL, "device index out of range");
    desv_batch b = new (std::nothrow) desv_batch_s();
    if (!b) return bfail(DESV_ENOMEM, "host allocation failed");
    b->dtype = dtype;
    b->device = device;
    *out = b;
    return DESV_OK;
}

int desv_batch_destroy(desv_batch b) {
    if (!b) return DESV_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->release();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return DESV_OK;
}

int desv_batch_add(desv_batch b, int n, const double *ops, int64_t n_ops, const int64_t *param_of_op,
                   int64_t n_params, const uint64_t *xmask, const uint64_t *zmask, const double *coeff,
                   int64_t n_terms, int64_t *index) {
    if (!b) return bfail(DESV_EINVAL, "null batch");
    if (n < 1 || n > 34 || n_ops < 0 || n_terms < 0 || n_params < 0) return bfail(DESV_EINVAL, "bad sizes");
    if ((n_ops && (!ops || !param_of_op)) || (n_terms && (!xmask || !zmask || !coeff)))
        return bfail(DESV_EINVAL, "null arrays");
    BInst in;
    in.n = n;
    in.rows.assign(ops, ops + 4 * n_ops);
    in.pmap.assign(param_of_op, param_of_op + n_ops);
    for (int64_t i = 0; i < n_ops; ++i)
        if (in.pmap[i] >= n_params) return bfail(DESV_EINVAL, "param_of_op index out of range");
    in.n_params = n_params;
    const uint64_t all = (1ull << n) - 1;
    for (int64_t t = 0; t < n_terms; ++t)
        if ((xmask[t] | zmask[t]) & ~all) return bfail(DESV_EINVAL, "Pauli term references a qubit out of range");
    in.x.assign(xmask, xmask + n_terms);
    in.z.assign(zmask, zmask + n_terms);
    in.coeff.assign(coeff, coeff + n_terms);
    // validate structure now (unknown gates, bad qubits, non-differentiable slots)
    if (int r = decode_instance(in, true)) return r;
    b->inst.push_back(std::move(in));
    b->built = false;
    if (index) *index = (int64_t)b->inst.size() - 1;
    return DESV_OK;
}

int desv_batch_set_params(desv_batch b, int64_t index, const double *theta) {
    if (!b || index < 0 || index >= (int64_t)b->inst.size()) return bfail(DESV_EINVAL, "bad instance index");
    BInst &in = b->inst[index];
    if (in.n_params && !theta) return bfail(DESV_EINVAL, "null theta");
    for (size_t i = 0; i < in.pmap.size(); ++i)
        if (in.pmap[i] >= 0) in.rows[4 * i + 3] = theta[in.pmap[i]];
    in.dirty = true;
    return DESV_OK;
}

int desv_batch_set_initial_state(desv_batch b, int64_t index, const void *host_amps) {
    if (!b || index < 0 || index >= (int64_t)b->inst.size()) return bfail(DESV_EINVAL, "bad instance index");
    BInst &in = b->inst[index];
    const size_t sb = ((size_t)1 << in.n) * abytes(b->dtype);
    const bool had = !in.init.empty();
    b->reverse_ready = false;
    if (!host_amps) {
        in.init.clear();
        if (had) b->built = false;
        return DESV_OK;
    }
    in.init.assign(reinterpret_cast<const unsigned char *>(host_amps),
                   reinterpret_cast<const unsigned char *>(host_amps) + sb);
    if (b->built && had) {
        BHIP(hipMemcpy(in.init_dev, in.init.data(), sb, hipMemcpyHostToDevice));
    } else {
        b->built = false;
    }
    return DESV_OK;
}

int desv_batch_configure(desv_batch b, int tile_bits, int fuse_bits, int use_mfma, int use_graph) {
    if (!b) return bfail(DESV_EINVAL, "null batch");
    if (tile_bits < 0 || tile_bits > MAX_TILE_BITS) return bfail(DESV_EINVAL, "tile_bits out of range");
    if (fuse_bits == 0) fuse_bits = 2;
    if (fuse_bits < 1 || fuse_bits > 5) return bfail(DESV_EINVAL, "fuse_bits must be in [1, 5]");
    const int mf = use_mfma != 0, ug = use_graph != 0;
    if (mf && b->dtype != DESV_C64)
        return bfail(DESV_EINVAL, "use_mfma applies to complex64 batches only (complex128 blocks run on VALU)");
    if (b->tile_bits != tile_bits || b->fuse_bits != fuse_bits || b->use_mfma != mf || b->use_graph != ug)
        b->built = false;
    b->tile_bits = tile_bits;
    b->fuse_bits = fuse_bits;
    b->use_mfma = mf;
    b->use_graph = ug;
    return DESV_OK;
}

int desv_batch_set_forward_only(desv_batch b, int enable) {
    if (!b) return bfail(DESV_EINVAL, "null batch");
    const int f = enable ? 1 : 0;
    if (f != b->forward_only) b->built = false;
    b->forward_only = f;
    return DESV_OK;
}

int desv_batch_get_states(desv_batch b, void *host) {
    if (!b || !host) return bfail(DESV_EINVAL, "null argument");
    if (!b->built || !b->states_dev) return bfail(DESV_EINVAL, "the batch has not run yet");
    BHIP(hipSetDevice(b->device));
    BHIP(hipMemcpyAsync(host, b->states_dev, b->states_bytes, hipMemcpyDeviceToHost, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return DESV_OK;
}

int desv_batch_set_compensated(desv_batch b, int enable) {
    if (!b) return bfail(DESV_EINVAL, "null batch");
    const int c = enable && b->dtype == DESV_C64 ? 1 : 0;
    if (c != b->compensated) b->built = false;
    b->compensated = c;
    return DESV_OK;
}

int desv_batch_run(desv_batch b, double *energies, double *grads) {
    if (!b) return bfail(DESV_EINVAL, "null batch");
    if (b->forward_only && (energies || grads))
        return bfail(DESV_EINVAL, "a forward-only batch has no energies or gradients: pass NULL for both");
    auto t0 = std::chrono::steady_clock::now();
    if (int r = execute(b, grads != nullptr ? 1 : 0)) return r;
    int64_t goff = 0;
    for (size_t i = 0; i < b->inst.size(); ++i) {
        const BInst &in = b->inst[i];
        if (energies) energies[i] = b->out_host[in.e_out];
        if (grads) {
            for (int64_t k = 0; k < in.n_params; ++k) grads[goff + k] = 0.0;
            for (size_t s = 0; s < in.slot_param.size(); ++s)
                grads[goff + in.slot_param[s]] += b->out_host[in.s_out + s];
        }
        goff += in.n_params;
    }
    b->last_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    return DESV_OK;
}

int desv_batch_run_reverse(desv_batch b, double *grads) {
    if (!b || !grads) return bfail(DESV_EINVAL, "null argument");
    if (b->forward_only) return bfail(DESV_EINVAL, "a forward-only batch has no gradients");
    auto t0 = std::chrono::steady_clock::now();
    if (int r = execute(b, 2)) return r;
    int64_t goff = 0;
    for (size_t i = 0; i < b->inst.size(); ++i) {
        const BInst &in = b->inst[i];
        for (int64_t k = 0; k < in.n_params; ++k) grads[goff + k] = 0.0;
        for (size_t s = 0; s < in.slot_param.size(); ++s) grads[goff + in.slot_param[s]] += b->out_host[in.s_out + s];
        goff += in.n_params;
    }
    b->last_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    return DESV_OK;
}

int desv_batch_state_device_ptr(desv_batch b, int64_t index, void **out) {
    if (!b || !out || index < 0 || index >= (int64_t)b->inst.size()) return bfail(DESV_EINVAL, "bad argument");
    if (!b->built) return bfail(DESV_EINVAL, "batch has not been run");
    *out = b->inst[index].psi;
    return DESV_OK;
}

int desv_batch_get_state(desv_batch b, int64_t index, void *host_amps) {
    if (!b || index < 0 || index >= (int64_t)b->inst.size()) return bfail(DESV_EINVAL, "bad instance index");
    if (!b->built) return bfail(DESV_EINVAL, "batch has not been run");
    if (!host_amps) return bfail(DESV_EINVAL, "null host buffer");
    const BInst &in = b->inst[index];
    BHIP(hipMemcpyAsync(host_amps, in.psi, ((size_t)1 << in.n) * abytes(b->dtype), hipMemcpyDeviceToHost,
                        b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    return DESV_OK;
}

int desv_batch_last_time_us(desv_batch b, double *us) {
    if (!b || !us) return bfail(DESV_EINVAL, "null argument");
    *us = b->last_us;
    return DESV_OK;
}

int desv_batch_completion_us(desv_batch b, double *us) {
    if (!b || !us) return bfail(DESV_EINVAL, "null argument");
    if (!b->built || !b->stamps_dev) return bfail(DESV_EINVAL, "the batch has not run yet");
    BHIP(hipSetDevice(b->device));
    std::vector<unsigned long long> st((size_t)b->n_stamps + 1);
    BHIP(hipMemcpyAsync(st.data(), b->stamps_dev, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                        b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    int khz = 0;
    BHIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
    if (khz <= 0) return bfail(DESV_EDEVICE, "device reports no wall clock rate");
    for (int64_t i = 0; i < b->n_stamps; ++i)
        us[i] = st[1 + i] >= st[0] ? (double)(st[1 + i] - st[0]) * 1e3 / khz : -1.0;
    return DESV_OK;
}

}  // extern "C"
