// Host-side circuit compiler: op-table decoding, gate fusion and
// cache-blocked stage scheduling for the LDS-tile engine.
//
// The reference applies one gate per full-vector pass (compiler.py:136-315,
// scanned at :333).  Here a circuit is cut into *stages*: each stage picks a
// set of m basis-index bits (always including the low bits, so global loads
// are contiguous runs), and every gate whose non-diagonal action stays inside
// that set is applied while a tile of 2^m amplitudes sits in LDS.  One stage
// costs one HBM read + write of the state no matter how many gates it holds.
// Diagonal gates and controls may involve bits outside the tile: those bits
// are constant over a tile.
#include "desv_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace desv {

namespace {

const double kInv2 = 0.70710678118654752440;

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

// 2x2 matrix of 1-qubit gate ids 0..14 (reference compiler.py:52-134).
void mat_1q(int g, double p, cd m[4]) {
    const cd I(0, 1);
    double c = std::cos(0.5 * p), s = std::sin(0.5 * p);
    cd e = std::exp(I * p);
    m[0] = 1; m[1] = 0; m[2] = 0; m[3] = 1;
    switch (g) {
        case 1: m[0] = kInv2; m[1] = kInv2; m[2] = kInv2; m[3] = -kInv2; break;
        case 2: m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; break;
        case 3: m[0] = 0; m[1] = -I; m[2] = I; m[3] = 0; break;
        case 4: m[3] = -1; break;
        case 5: m[3] = I; break;
        case 6: m[3] = -I; break;
        case 7: m[3] = std::exp(I * (M_PI / 4)); break;
        case 8: m[3] = std::exp(-I * (M_PI / 4)); break;
        case 9: m[0] = c; m[1] = -I * s; m[2] = -I * s; m[3] = c; break;
        case 10: m[0] = c; m[1] = -s; m[2] = s; m[3] = c; break;
        case 11: m[0] = std::exp(-0.5 * I * p); m[3] = std::exp(0.5 * I * p); break;
        case 12: m[3] = e; break;
        case 13: m[0] = cd(0.5, 0.5); m[1] = cd(0.5, -0.5); m[2] = cd(0.5, -0.5); m[3] = cd(0.5, 0.5); break;
        case 14: m[0] = e; m[3] = e; break;
        default: break;
    }
}

bool is_diag_1q(int g) {
    return g == 0 || g == 4 || g == 5 || g == 6 || g == 7 || g == 8 || g == 11 || g == 12 || g == 14;
}

}  // namespace

bool decode_row(const double *row, int n, std::vector<Gate> &out, std::string &err) {
    double gd = row[0];
    int g = (int)gd;
    if ((double)g != gd || !((g >= 0 && g <= 14) || (g >= 20 && g <= 25))) {
        err = "unknown gate id " + std::to_string(gd);
        return false;
    }
    double fq1 = row[1], fq2 = row[2], p = row[3];
    int q1 = (int)fq1;
    if ((double)q1 != fq1 || q1 < 0 || q1 >= n) {
        err = "Qubit index " + std::to_string((long long)fq1) + " out of range [0, " + std::to_string(n) + ")";
        return false;
    }
    const cd I(0, 1);
    if (g <= 14) {
        if (g == 0) return true;
        cd m[4];
        mat_1q(g, p, m);
        Gate gt;
        gt.k = 1;
        gt.bits[0] = n - 1 - q1;
        if (g == 14) {
            gt.kind = G_SCALAR;
            gt.k = 0;
            gt.m = {m[0]};
        } else if (is_diag_1q(g)) {
            gt.kind = G_DIAG;
            gt.m = {m[0], m[3]};
        } else {
            gt.kind = G_DENSE;
            gt.m.assign(m, m + 4);
        }
        out.push_back(gt);
        return true;
    }
    int q2 = (int)fq2;
    if ((double)q2 != fq2 || q2 < 0 || q2 >= n) {
        err = "Qubit index " + std::to_string((long long)fq2) + " out of range [0, " + std::to_string(n) + ")";
        return false;
    }
    if (q1 == q2) {
        err = "Control and target qubits must differ";
        return false;
    }
    int cb = n - 1 - q1, tb = n - 1 - q2;
    Gate gt;
    gt.k = 1;
    gt.bits[0] = tb;
    gt.cmask = 1ull << cb;
    switch (g) {
        case 20: gt.kind = G_DENSE; gt.m = {0, 1, 1, 0}; break;
        case 21: gt.kind = G_DIAG; gt.m = {1, -1}; break;
        case 22: gt.kind = G_DIAG; gt.m = {1, std::exp(I * p)}; break;
        case 23:
            gt.kind = G_DENSE;
            gt.k = 2;
            gt.cmask = 0;
            gt.bits[0] = tb;
            gt.bits[1] = cb;
            gt.m.assign(16, 0);
            gt.m[0] = 1; gt.m[1 * 4 + 2] = 1; gt.m[2 * 4 + 1] = 1; gt.m[15] = 1;
            break;
        case 24: gt.kind = G_DENSE; gt.m = {0, -I, I, 0}; break;
        case 25: gt.kind = G_DIAG; gt.m = {std::exp(-0.5 * I * p), std::exp(0.5 * I * p)}; break;
        default: err = "unknown gate id " + std::to_string(g); return false;
    }
    out.push_back(gt);
    return true;
}

bool decode_matrix(const double *mat, const int *qubits, int k, int n, Gate &g, std::string &err) {
    if (k < 1 || k > 5) {
        err = "matrix gates act on 1..5 qubits";
        return false;
    }
    uint64_t seen = 0;
    for (int j = 0; j < k; ++j) {
        if (qubits[j] < 0 || qubits[j] >= n) {
            err = "Qubit index " + std::to_string(qubits[j]) + " out of range [0, " + std::to_string(n) + ")";
            return false;
        }
        uint64_t b = 1ull << (n - 1 - qubits[j]);
        if (seen & b) {
            err = "Control and target qubits must differ";
            return false;
        }
        seen |= b;
    }
    g = Gate();
    g.kind = G_DENSE;
    g.k = k;
    for (int i = 0; i < k; ++i) g.bits[i] = n - 1 - qubits[k - 1 - i];
    int d = 1 << k;
    g.m.resize((size_t)d * d);
    for (int i = 0; i < d * d; ++i) g.m[i] = cd(mat[2 * i], mat[2 * i + 1]);
    return true;
}

Gate inverse_gate(const Gate &g) {
    Gate r = g;
    if (g.kind == G_DENSE) {
        int d = 1 << g.k;
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) r.m[(size_t)i * d + j] = std::conj(g.m[(size_t)j * d + i]);
    } else {
        for (auto &x : r.m) x = std::conj(x);
    }
    return r;
}

bool grad_gate_for_row(const double *row, int n, int slot, Gate &g, std::string &err) {
    int gid = (int)row[0];
    int q1 = (int)row[1], q2 = (int)row[2];
    const cd I(0, 1);
    g = Gate();
    g.kind = G_GRAD;
    g.k = 1;
    g.grad_slot = slot;
    g.bits[0] = n - 1 - q1;
    // M = 2 * G with dU/dtheta = G U, so grad += Re<lam| M |psi>.
    switch (gid) {
        case 9: g.grad_diag = false; g.m = {0, -I, -I, 0}; break;         // -i X
        case 10: g.grad_diag = false; g.m = {0, -1, 1, 0}; break;         // -i Y
        case 11: g.grad_diag = true; g.m = {-I, I}; break;                // -i Z
        case 12: g.grad_diag = true; g.m = {0, 2.0 * I}; break;           // 2i |1><1|
        case 22:
            g.grad_diag = true; g.m = {0, 2.0 * I};
            g.bits[0] = n - 1 - q2; g.cmask = 1ull << (n - 1 - q1); break;
        case 25:
            g.grad_diag = true; g.m = {-I, I};
            g.bits[0] = n - 1 - q2; g.cmask = 1ull << (n - 1 - q1); break;
        default:
            err = "gate id " + std::to_string(gid) + " does not carry a differentiable parameter";
            return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Fusion
// ---------------------------------------------------------------------------
namespace {

// Apply gate g (global bits) to every column of the dense block matrix B that
// acts on `ub` (LSB-first list of K global bits): B <- G B.
void left_apply(std::vector<cd> &B, const std::vector<int> &ub, const Gate &g) {
    int K = (int)ub.size();
    int D = 1 << K;
    auto local = [&](int b) {
        for (int i = 0; i < K; ++i)
            if (ub[i] == b) return i;
        return -1;
    };
    uint32_t cl = 0;
    for (int b = 0; b < 64; ++b)
        if (g.cmask >> b & 1) cl |= 1u << local(b);
    if (g.kind == G_SCALAR) {
        for (auto &x : B) x *= g.m[0];
        return;
    }
    int lp[5];
    for (int j = 0; j < g.k; ++j) lp[j] = local(g.bits[j]);
    uint32_t tm = 0;
    for (int j = 0; j < g.k; ++j) tm |= 1u << lp[j];
    int gd = 1 << g.k;
    std::vector<cd> tmp(gd);
    for (int col = 0; col < D; ++col) {
        for (int r0 = 0; r0 < D; ++r0) {
            if (r0 & tm) continue;
            if ((r0 & cl) != cl) continue;
            int idx[32];
            for (int s = 0; s < gd; ++s) {
                int r = r0;
                for (int j = 0; j < g.k; ++j)
                    if (s >> j & 1) r |= 1 << lp[j];
                idx[s] = r;
            }
            if (g.kind == G_DIAG) {
                for (int s = 0; s < gd; ++s) B[(size_t)idx[s] * D + col] *= g.m[s];
            } else {
                for (int s = 0; s < gd; ++s) {
                    cd acc = 0;
                    for (int t = 0; t < gd; ++t) acc += g.m[(size_t)s * gd + t] * B[(size_t)idx[t] * D + col];
                    tmp[s] = acc;
                }
                for (int s = 0; s < gd; ++s) B[(size_t)idx[s] * D + col] = tmp[s];
            }
        }
    }
}

struct Block {
    uint64_t bits = 0;
    std::vector<int> gates;  // indices into the stage gate list, in order
};

// Turn a block into one gate: a lone gate stays as it is; otherwise the
// product is formed over the union bits (all-diagonal products stay diagonal).
Gate materialise(const Block &b, const std::vector<Gate> &gs) {
    if (b.gates.size() == 1) return gs[b.gates[0]];
    std::vector<int> ub;
    for (int i = 0; i < 64; ++i)
        if (b.bits >> i & 1) ub.push_back(i);
    int K = (int)ub.size();
    int D = 1 << K;
    bool all_diag = true;
    for (int gi : b.gates) all_diag &= gs[gi].kind == G_DIAG || gs[gi].kind == G_SCALAR;
    Gate out;
    out.k = K;
    for (int i = 0; i < K; ++i) out.bits[i] = ub[i];
    out.src = gs[b.gates[0]].src;
    if (all_diag) {
        out.kind = G_DIAG;
        out.m.assign(D, 1.0);
        for (int gi : b.gates) {
            const Gate &g = gs[gi];
            for (int s = 0; s < D; ++s) {
                bool on = true;
                int sub = 0;
                for (int i = 0; i < K; ++i) {
                    bool v = s >> i & 1;
                    if ((g.cmask >> ub[i] & 1) && !v) on = false;
                    for (int j = 0; j < g.k; ++j)
                        if (g.kind == G_DIAG && g.bits[j] == ub[i] && v) sub |= 1 << j;
                }
                if (!on) continue;
                out.m[s] *= (g.kind == G_SCALAR) ? g.m[0] : g.m[sub];
            }
        }
        return out;
    }
    out.kind = G_DENSE;
    out.m.assign((size_t)D * D, 0.0);
    for (int i = 0; i < D; ++i) out.m[(size_t)i * D + i] = 1.0;
    for (int gi : b.gates) left_apply(out.m, ub, gs[gi]);
    return out;
}

// Greedy fusion of a gate sequence into blocks of at most K bits.  A gate is
// fusable when it is not a gradient probe, and all its bits lie in `allowed`.
// Open blocks are always the latest writer of their bits, so merging open
// blocks that touch a new gate keeps a valid order.
std::vector<Gate> fuse(const std::vector<Gate> &gs, int K, uint64_t allowed, std::vector<Recipe> *rec) {
    std::vector<Gate> out;
    std::vector<Block> open;
    auto emit = [&](const Block &b) {
        out.push_back(materialise(b, gs));
        if (rec) {
            Recipe r;
            r.gates = b.gates;
            r.bits = b.gates.size() == 1 ? 0 : b.bits;
            rec->push_back(r);
        }
    };
    auto emit_touching = [&](uint64_t mask) {
        std::vector<Block> keep;
        for (auto &b : open) {
            if (b.bits & mask) emit(b);
            else keep.push_back(std::move(b));
        }
        open.swap(keep);
    };
    for (int i = 0; i < (int)gs.size(); ++i) {
        const Gate &g = gs[i];
        uint64_t am = g.all_mask();
        bool fusable = g.kind != G_GRAD && (am & ~allowed) == 0 && popc(am) <= K;
        if (g.kind == G_SCALAR) fusable = false;
        if (!fusable) {
            emit_touching(am);
            out.push_back(g);
            if (rec) rec->push_back(Recipe{{i}, 0});
            continue;
        }
        uint64_t u = am;
        for (auto &b : open)
            if (b.bits & am) u |= b.bits;
        if (popc(u) <= K) {
            Block nb;
            nb.bits = u;
            std::vector<Block> keep;
            for (auto &b : open) {
                if (b.bits & am) nb.gates.insert(nb.gates.end(), b.gates.begin(), b.gates.end());
                else keep.push_back(std::move(b));
            }
            nb.gates.push_back(i);
            keep.push_back(std::move(nb));
            open.swap(keep);
        } else {
            // Too wide for one block: emit the widest touching blocks until
            // the rest fit with the gate, and merge the rest into it.  The
            // touching blocks are disjoint and all precede the gate, so any
            // of them may go first (a lone 1-qubit gate waiting on the new
            // gate's other bit is absorbed instead of emitted on its own).
            std::vector<int> touch;
            for (int b = 0; b < (int)open.size(); ++b)
                if (open[b].bits & am) touch.push_back(b);
            std::stable_sort(touch.begin(), touch.end(),
                             [&](int x, int y) { return popc(open[x].bits) < popc(open[y].bits); });
            uint64_t keep_bits = am;
            std::vector<char> merge(open.size(), 0);
            for (int b : touch)
                if (popc(keep_bits | open[b].bits) <= K) {
                    keep_bits |= open[b].bits;
                    merge[b] = 1;
                }
            Block nb;
            nb.bits = keep_bits;
            std::vector<Block> keep;
            for (int b = 0; b < (int)open.size(); ++b) {
                if (merge[b]) nb.gates.insert(nb.gates.end(), open[b].gates.begin(), open[b].gates.end());
                else if (open[b].bits & am) emit(open[b]);
                else keep.push_back(std::move(open[b]));
            }
            nb.gates.push_back(i);
            keep.push_back(std::move(nb));
            open.swap(keep);
        }
    }
    for (auto &b : open) emit(b);
    return out;
}

uint64_t low_mask(int L) { return L >= 64 ? ~0ull : ((1ull << L) - 1); }

// Fill a tile mask up to m bits with the lowest free bits.
uint64_t fill_tile(uint64_t tile, int m, int n) {
    for (int b = 0; b < n && popc(tile) < m; ++b) tile |= 1ull << b;
    return tile;
}

const int kMaxGradPerStage = 64;

bool lower_stage(const std::vector<Gate> &ops, uint64_t tile, int n, const CompileOpts &o,
                 Plan &plan, std::string &err, const int *sperm = nullptr) {
    DevStage st;
    std::memset(&st, 0, sizeof(st));
    st.m = popc(tile);
    if (st.m > MAX_TILE_BITS) {
        err = "tile wider than the engine supports";
        return false;
    }
    for (int j = 0; j < st.m; ++j) {
        st.sperm[j] = sperm ? sperm[j] : j;
        if (st.sperm[j] != j) st.permuted = 1;
    }
    int L = 0;
    while (L < n && (tile >> L & 1)) ++L;
    st.L = L;
    st.nh = 0;
    for (int b = L; b < n; ++b)
        if (tile >> b & 1) {
            if (st.nh >= MAX_HBITS) {
                err = "tile has too many high bits";
                return false;
            }
            st.hbits[st.nh++] = b;
        }
    st.tile_mask = tile;
    st.n_tiles = 1ll << (n - st.m);
    st.op_begin = (int)plan.ops.size();
    int lpos[64];
    for (int b = 0, c = 0; b < 64; ++b) lpos[b] = (b < n && (tile >> b & 1)) ? c++ : -1;
    int ngrad = 0;
    for (const Gate &g : ops) {
        DevOp d;
        std::memset(&d, 0, sizeof(d));
        d.lt = -1;
        d.k = g.k;
        d.mat = (int)plan.mats.size();
        d.slot = g.grad_slot;
        for (int b = 0; b < n; ++b)
            if (g.cmask >> b & 1) {
                if (lpos[b] >= 0) d.cloc |= 1u << lpos[b];
                else d.cglob |= 1ull << b;
            }
        for (int j = 0; j < g.k; ++j) {
            int b = g.bits[j];
            d.pos[j] = lpos[b] >= 0 ? lpos[b] : -1 - b;
        }
        bool need_local = g.kind == G_DENSE || (g.kind == G_GRAD && !g.grad_diag);
        if (need_local)
            for (int j = 0; j < g.k; ++j)
                if (d.pos[j] < 0) {
                    err = "internal: non-diagonal target outside the tile";
                    return false;
                }
        switch (g.kind) {
            case G_SCALAR: d.kind = OPK_SCALAR; break;
            case G_DIAG: d.kind = OPK_DIAG; break;
            case G_DENSE:
                if (g.k == 1) d.kind = OPK_DENSE1;
                else if (g.k == 2) d.kind = OPK_DENSE2;
                else if (o.use_mfma && g.k >= 3) { d.kind = OPK_MFMA; plan.mfma_blocks++; }
                else d.kind = OPK_DENSEK;
                break;
            case G_GRAD:
                if (g.grad_diag) d.kind = OPK_GRADDIAG;
                else d.kind = OPK_GRAD;
                if (d.kind == OPK_GRADDIAG)
                    for (int j = 0; j < g.k; ++j)
                        if (d.pos[j] < 0) {
                            err = "internal: gradient probe outside the tile";
                            return false;
                        }
                ++ngrad;
                break;
        }
        plan.mats.insert(plan.mats.end(), g.m.begin(), g.m.end());
        plan.ops.push_back(d);
    }
    st.op_end = (int)plan.ops.size();
    st.n_grad = ngrad;
    // Reverse sweeps: a gradient probe followed by a 1-/2-qubit dense or
    // diagonal block that contains the probe's bit (the inverse of the
    // probed gate, possibly fused with its neighbours) run as one sweep of
    // the tile; pos[4] = 1 on the probe marks the pair.
    if (o.nv == 2)
        for (int i = st.op_begin; i + 1 < st.op_end; ++i) {
            DevOp &gp = plan.ops[i];
            const DevOp &nx = plan.ops[i + 1];
            if ((gp.kind != OPK_GRAD && gp.kind != OPK_GRADDIAG) || gp.k != 1 || gp.cloc || gp.cglob ||
                gp.pos[0] < 0)
                continue;
            if ((nx.kind != OPK_DENSE1 && nx.kind != OPK_DENSE2 && nx.kind != OPK_DIAG) || nx.k < 1 || nx.k > 2 ||
                nx.cloc || nx.cglob)
                continue;
            bool local = true, has = false;
            for (int j = 0; j < nx.k; ++j) {
                local &= nx.pos[j] >= 0;
                has |= nx.pos[j] == gp.pos[0];
            }
            if (!local || !has) continue;
            gp.pos[4] = 1;
            ++i;
        }
    // kernel variant: 3 if any 5-qubit MFMA block (its fragments double the
    // register budget), 2 if any MFMA block (then no k >= 4 VALU blocks), 1
    // if any k >= 3 VALU block, else 0
    st.heavy = 0;
    for (int i = st.op_begin; i < st.op_end; ++i) {
        if (plan.ops[i].kind == OPK_MFMA) st.heavy = std::max(st.heavy, plan.ops[i].k == 5 ? 3 : 2);
        else if (plan.ops[i].kind == OPK_DENSEK && st.heavy == 0) st.heavy = 1;
    }
    plan.stages.push_back(st);
    return true;
}

}  // namespace

namespace {

// Scan up to `window` remaining gates with tile (logical bit mask) S: a gate
// is taken when its non-diagonal bits lie in S and none of its bits is
// blocked; a gate that cannot be taken blocks all of its bits, so taken
// gates commute with everything deferred.  Returns the number taken.
int scan_take(const std::vector<Gate> &gates, const std::vector<int> &rem, uint64_t S, uint64_t all, int window,
              std::vector<int> *take, std::vector<int> *defer) {
    uint64_t blocked = 0;
    int cnt = 0, ngrad = 0;
    const int lim = (int)std::min<size_t>(rem.size(), (size_t)window);
    int r = 0;
    for (; r < lim; ++r) {
        if (blocked == all) break;
        const int gi = rem[r];
        const Gate &g = gates[gi];
        const uint64_t am = g.all_mask();
        const bool grad_full = g.kind == G_GRAD && ngrad >= kMaxGradPerStage;
        if ((am & blocked) || (g.tile_mask() & ~S) || grad_full) {
            blocked |= am;
            if (defer) defer->push_back(gi);
            continue;
        }
        if (g.kind == G_GRAD) ++ngrad;
        ++cnt;
        if (take) take->push_back(gi);
    }
    if (defer)
        for (; r < (int)rem.size(); ++r) defer->push_back(rem[r]);
    return cnt;
}

// A stage's trailing 1-qubit gates (dense or diagonal) on a qubit that has a
// later gate among the deferred ones move to the deferred list: they commute
// with everything off their qubit, and in the later stage they fuse into the
// block of that gate instead of costing a tile sweep of their own here (the
// last gates of a qubit before a stage boundary are otherwise emitted as lone
// 1-qubit ops).  Per-qubit order is kept: only a qubit's final run in `take`
// moves.  Both lists stay in gate order; `take` is never emptied.
void defer_trailing_singles(const std::vector<Gate> &gates, std::vector<int> &take, std::vector<int> &defer) {
    uint64_t later = 0;
    for (int gi : defer) later |= gates[gi].all_mask();
    uint64_t closed = 0;
    std::vector<char> mv(take.size(), 0);
    size_t nmv = 0;
    for (int t = (int)take.size() - 1; t >= 0; --t) {
        const Gate &g = gates[take[t]];
        const uint64_t am = g.all_mask();
        if (popc(am) == 1 && (g.kind == G_DENSE || g.kind == G_DIAG) && (later & am) && !(closed & am)) {
            mv[t] = 1;
            ++nmv;
        } else {
            closed |= am;
        }
    }
    if (nmv == 0 || nmv == take.size()) return;
    std::vector<int> t2, d2;
    t2.reserve(take.size() - nmv);
    d2.reserve(defer.size() + nmv);
    size_t j = 0;
    for (size_t t = 0; t < take.size(); ++t) {
        if (!mv[t]) {
            t2.push_back(take[t]);
            continue;
        }
        while (j < defer.size() && defer[j] < take[t]) d2.push_back(defer[j++]);
        d2.push_back(take[t]);
    }
    while (j < defer.size()) d2.push_back(defer[j++]);
    take.swap(t2);
    defer.swap(d2);
}

// Reverse sweeps apply every op to both psi and lambda, so a gradient probe
// Re<lam|G_q|psi> commutes with ops on other bits: each uncontrolled 1-qubit
// probe moves forward to just before the first later op that touches its
// bit, when that op is an uncontrolled 1-/2-qubit dense or diagonal block (the
// kernel then runs the two in one sweep).  A probe separated from that op by
// probes only stays (two probes before one 2-qubit block would otherwise
// displace each other forever); every move removes non-probe ops from between
// a probe and its block, so the loop ends.  Recipes move with their ops.
void move_probes_to_blocks(std::vector<Gate> &g, std::vector<Recipe> &rec) {
    for (size_t i = 0; i < g.size(); ++i) {
        if (g[i].kind != G_GRAD || g[i].cmask || g[i].k != 1) continue;
        const uint64_t pb = 1ull << g[i].bits[0];
        size_t j = i + 1;
        bool only_probes = true;
        while (j < g.size() && !(g[j].all_mask() & pb)) {
            only_probes &= g[j].kind == G_GRAD;
            ++j;
        }
        if (j == g.size() || only_probes) continue;
        const Gate &b = g[j];
        if (!((b.kind == G_DENSE || b.kind == G_DIAG) && b.k >= 1 && b.k <= 2 && b.cmask == 0)) continue;
        std::rotate(g.begin() + i, g.begin() + i + 1, g.begin() + j);
        std::rotate(rec.begin() + i, rec.begin() + i + 1, rec.begin() + j);
        --i;  // position i now holds the next op
    }
}

Gate to_physical(const Gate &g, const std::vector<int> &pi) {
    Gate p = g;
    for (int j = 0; j < g.k; ++j) p.bits[j] = pi[g.bits[j]];
    p.cmask = 0;
    for (int b = 0; b < 64; ++b)
        if (g.cmask >> b & 1) p.cmask |= 1ull << pi[b];
    return p;
}

// Store permutation of a stage whose tile holds logical bits S (physical
// bits P) when the logical->physical map changes from pi to pi2 (pi2 only
// moves bits of S within P).  sperm[j] = load-local bit that ends up in the
// j-th lowest physical bit of P.
void store_perm(uint64_t P, uint64_t S, const std::vector<int> &pi, const std::vector<int> &pi2, int n,
                int *sperm) {
    int lpos[64];
    for (int b = 0, c = 0; b < 64; ++b) lpos[b] = (b < n && (P >> b & 1)) ? c++ : -1;
    int j = 0;
    for (int b = 0; b < n; ++b) {
        if (!(P >> b & 1)) continue;
        int src = -1;
        for (int u = 0; u < n; ++u)
            if ((S >> u & 1) && pi2[u] == b) src = lpos[pi[u]];
        sperm[j++] = src;
    }
}

// Relabelling store of a stage whose tile is P under layout pi (logical u at
// physical pi[u]) into the identity layout: store-order bit j (the j-th lowest
// logical bit of the tile) takes tile-local bit sperm[j]; the i-th lowest
// physical bit outside P goes to the logical index of its occupant.
void identity_store(uint64_t P, const std::vector<int> &pi, int n, DevStage &st, StageRelabel &r) {
    std::vector<int> inv(n), lpos(n, -1);
    for (int u = 0; u < n; ++u) inv[pi[u]] = u;
    for (int b = 0, c = 0; b < n; ++b)
        if (P >> b & 1) lpos[b] = c++;
    r.store_mask = 0;
    for (int b = 0; b < n; ++b)
        if (P >> b & 1) r.store_mask |= 1ull << inv[b];
    int j = 0;
    for (int u = 0; u < n; ++u)
        if (r.store_mask >> u & 1) st.sperm[j++] = lpos[pi[u]];
    st.permuted = 1;
    r.omap.clear();
    for (int b = 0; b < n; ++b)
        if (!(P >> b & 1)) r.omap.push_back(inv[b]);
}

// Replace the restore passes (stages n_compute ..) by relabelling stages: the
// last compute stage itself when its tile holds the qubits of the low bits
// 0 .. L-1, else one restore pass whose tile holds the low bits' occupants
// and those qubits.  The in-place stages are kept as the fallback.
void relabel_tail(Plan &plan, int n, int m, int L, const CompileOpts &o, const std::vector<int> &last_pi,
                  uint64_t last_S, const std::vector<int> &end_pi, size_t n_compute) {
    const size_t n_restore = plan.stages.size() - n_compute;
    const bool last_has_low = (last_S & low_mask(L)) == low_mask(L);
    if (!last_has_low && n_restore < 2) return;  // one relabelling pass would save nothing
    const size_t from = last_has_low ? n_compute - 1 : n_compute;
    plan.fb_from = (int)from;
    plan.fb_stages.assign(plan.stages.begin() + (long)from, plan.stages.end());
    plan.stages.resize(from);
    plan.relabel.assign(from, StageRelabel());
    StageRelabel r;
    if (last_has_low) {
        DevStage st = plan.fb_stages[0];
        identity_store(st.tile_mask, last_pi, n, st, r);
        plan.stages.push_back(st);
    } else {
        uint64_t P = low_mask(L);
        for (int u = 0; u < L; ++u) P |= 1ull << end_pi[u];
        P = fill_tile(P, m, n);
        std::string err;
        if (!lower_stage({}, P, n, o, plan, err)) {
            // cannot happen for a tile the plan's other stages already use;
            // keep the in-place tail
            plan.stages.insert(plan.stages.end(), plan.fb_stages.begin(), plan.fb_stages.end());
            plan.fb_stages.clear();
            plan.relabel.clear();
            plan.fb_from = -1;
            return;
        }
        identity_store(P, end_pi, n, plan.stages.back(), r);
    }
    plan.relabel.push_back(r);
}

}  // namespace

// Score perturbation of the randomised planner passes: a splitmix64 stream,
// uniform in [0, 1) (null: the deterministic greedy).
struct PlanNoise {
    uint64_t state;
    double next() {
        uint64_t z = (state += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

// Estimated cost of a plan in units of one block's work: a stage's load and
// store cost about two blocks, a pass with no blocks (layout restore) about
// five and a half (measured on the 30-qubit brickwork: 1.1 ms per stage plus
// 0.575 ms per block; 3.3 ms per restore pass).
static double plan_cost(const Plan &p) {
    double c = 0;
    for (const DevStage &st : p.stages) c += st.op_end > st.op_begin ? 2.0 + (st.op_end - st.op_begin) : 5.7;
    return c;
}

static bool build_plan_pass(const std::vector<Gate> &gates, int n, const CompileOpts &o, Plan &plan,
                            std::string &err, PlanNoise *noise);

bool build_plan(const std::vector<Gate> &gates, int n, const CompileOpts &o, Plan &plan,
                std::string &err) {
    if (!build_plan_pass(gates, n, o, plan, err, nullptr)) return false;
    if (o.search <= 0 || o.per_gate || plan.stages.size() < 3) return true;
    // Randomised restarts: every tile-growth choice scaled by a factor in
    // [1, 1.2); the cheapest plan by plan_cost wins (the deterministic one on
    // ties).  Seeds are fixed, so a program always compiles to the same plan.
    double best = plan_cost(plan);
    for (int s = 1; s <= o.search; ++s) {
        PlanNoise nz{(uint64_t)s * 0x2545f4914f6cdd1dull};
        Plan trial;
        std::string e2;
        if (!build_plan_pass(gates, n, o, trial, e2, &nz)) continue;
        const double c = plan_cost(trial);
        if (c < best - 1e-9) {
            best = c;
            plan = std::move(trial);
        }
    }
    return true;
}

static bool build_plan_pass(const std::vector<Gate> &gates, int n, const CompileOpts &o, Plan &plan,
                            std::string &err, PlanNoise *noise) {
    plan = Plan();
    plan.n = n;
    for (const Gate &g : gates)
        if (g.kind == G_GRAD) plan.n_grad_slots = std::max(plan.n_grad_slots, g.grad_slot + 1);
    int K = std::max(1, std::min(5, o.fuse_bits));
    int m = std::min(n, std::max(o.tile_bits, 1));
    int L = std::min(std::max(o.low_bits, 1), m);
    const uint64_t all = low_mask(n);
    if (o.per_gate && m < n && m < K + L) {
        err = "tile_bits too small for the fusion width";
        return false;
    }
    for (const Gate &g : gates)
        if (popc(g.tile_mask() | low_mask(L)) > m) {
            err = "gate wider than a tile";
            return false;
        }

    if (o.per_gate) {
        // One HBM pass per fused block.
        std::vector<Recipe> rec;
        std::vector<Gate> fused = fuse(gates, K, all, &rec);
        for (size_t i = 0; i < fused.size(); ++i) {
            const Gate &g = fused[i];
            uint64_t tile = fill_tile(low_mask(L) | g.tile_mask(), m, n);
            if (!lower_stage({g}, tile, n, o, plan, err)) return false;
            plan.recipes.push_back(rec[i]);
        }
        return true;
    }

    // Cache-blocked scheduling over a logical->physical bit map pi.
    std::vector<int> pi(n), inv(n);
    for (int u = 0; u < n; ++u) pi[u] = inv[u] = u;
    std::vector<int> rem(gates.size());
    for (size_t i = 0; i < gates.size(); ++i) rem[i] = (int)i;
    const int window = std::max(64, o.window);
    std::vector<int> last_pi;  // layout before the last compute stage
    uint64_t last_S = 0;       // its tile (logical bits)
    while (!rem.empty()) {
        uint64_t S = 0;  // logical bits in the tile
        for (int p = 0; p < L; ++p) S |= 1ull << inv[p];
        if (m < n) {
            // closure-greedy growth of the tile
            std::vector<int> def;
            while (popc(S) < m) {
                def.clear();
                const int base = scan_take(gates, rem, S, all, window, nullptr, &def);
                std::vector<uint64_t> cands;
                for (size_t i = 0; i < def.size() && i < 64; ++i) {
                    const uint64_t need = gates[def[i]].tile_mask() & ~S;
                    if (!need || popc(S | need) > m) continue;
                    if (std::find(cands.begin(), cands.end(), need) == cands.end()) cands.push_back(need);
                }
                if (cands.empty()) break;
                double best = 0;
                uint64_t pick = 0;
                for (uint64_t c : cands) {
                    const int t = scan_take(gates, rem, S | c, all, window, nullptr, nullptr);
                    double score = double(t - base) / popc(c);
                    if (noise) score *= 1.0 + 0.2 * noise->next();
                    if (score > best) {
                        best = score;
                        pick = c;
                    }
                }
                if (!pick) break;
                S |= pick;
            }
        } else {
            S = all;
        }
        std::vector<int> take, defer;
        scan_take(gates, rem, S, all, window, &take, &defer);
        if (take.empty()) {
            err = "internal: scheduler made no progress";
            return false;
        }
        if (m < n && K > 1) defer_trailing_singles(gates, take, defer);
        // Last compute stage: grow the tile from the bits its gates need and
        // the low bits' occupants by the occupant of each member's home bit
        // (so that member can go home), then by bits already home; the
        // write-back below then sends every bit it can home (the low bits
        // are not held for later gates), leaving the restore passes less.
        const bool endgame = defer.empty() && m < n && o.remap;
        if (endgame) {
            uint64_t T = 0;
            for (int gi : take) T |= gates[gi].tile_mask();
            for (int p = 0; p < L; ++p) T |= 1ull << inv[p];
            // a relabelling last stage needs the qubits of the low bits
            // (0 .. L-1) in its tile to write the identity layout coalesced
            if (o.relabel)
                for (int u = 0; u < L && popc(T) < m; ++u) T |= 1ull << u;
            if (popc(T) <= m) {
                for (bool grew = true; grew && popc(T) < m;) {
                    grew = false;
                    for (int u = 0; u < n && popc(T) < m; ++u)
                        if ((T >> u & 1) && pi[u] != u && !(T >> inv[u] & 1)) {
                            T |= 1ull << inv[u];
                            grew = true;
                        }
                }
                for (int u = 0; u < n && popc(T) < m; ++u)
                    if (pi[u] == u) T |= 1ull << u;
                S = T;  // holds every bit the stage's gates need
            }
        }
        // fill the tile: bits the next deferred gates need, then the lowest physical bits
        for (size_t i = 0; i < defer.size() && popc(S) < m && i < 256; ++i) {
            const uint64_t need = gates[defer[i]].tile_mask() & ~S;
            if (popc(S | need) <= m) S |= need;
        }
        for (int p = 0; p < n && popc(S) < m; ++p) S |= 1ull << inv[p];
        uint64_t P = 0;
        for (int u = 0; u < n; ++u)
            if (S >> u & 1) P |= 1ull << pi[u];
        // Next layout: the tile bits the deferred gates need soonest go to
        // the low physical bits; every other tile bit goes home if its home
        // bit is in the tile and free, else stays, else takes a free bit.
        std::vector<int> pi2 = pi;
        if (o.remap && m < n) {
            std::vector<int> want;
            for (size_t i = 0; i < defer.size() && (int)want.size() < L && i < 256; ++i) {
                const uint64_t tm = gates[defer[i]].tile_mask() & S;
                for (int u = 0; u < n; ++u)
                    if ((tm >> u & 1) && std::find(want.begin(), want.end(), u) == want.end() &&
                        (int)want.size() < L)
                        want.push_back(u);
            }
            for (int p = 0; p < L && (int)want.size() < L && !endgame; ++p)
                if (std::find(want.begin(), want.end(), inv[p]) == want.end()) want.push_back(inv[p]);
            uint64_t used = 0, placed = 0;
            for (int u : want)
                if (pi[u] < L) {
                    pi2[u] = pi[u];
                    used |= 1ull << pi[u];
                    placed |= 1ull << u;
                }
            for (int u : want)
                if (!(placed >> u & 1))
                    for (int p = 0; p < L; ++p)
                        if (!(used >> p & 1)) {
                            pi2[u] = p;
                            used |= 1ull << p;
                            placed |= 1ull << u;
                            break;
                        }
            for (int pass = 0; pass < 3; ++pass)
                for (int u = 0; u < n; ++u) {
                    if (!(S >> u & 1) || (placed >> u & 1)) continue;
                    int slot = -1;
                    if (pass == 0 && (P >> u & 1) && !(used >> u & 1) && (u >= L || endgame)) slot = u;
                    if (pass == 1 && !(used >> pi[u] & 1)) slot = pi[u];
                    if (pass == 2)
                        for (int b = 0; b < n; ++b)
                            if ((P >> b & 1) && !(used >> b & 1)) {
                                slot = b;
                                break;
                            }
                    if (slot < 0) continue;
                    pi2[u] = slot;
                    used |= 1ull << slot;
                    placed |= 1ull << u;
                }
        }
        std::vector<Gate> lg;
        lg.reserve(take.size());
        for (int gi : take) lg.push_back(gates[gi]);
        std::vector<Recipe> rec;
        std::vector<Gate> fused;
        if (K > 1) {
            fused = fuse(lg, K, S, &rec);
        } else {
            fused = lg;
            for (size_t i = 0; i < lg.size(); ++i) rec.push_back(Recipe{{(int)i}, 0});
        }
        if (o.nv == 2) move_probes_to_blocks(fused, rec);
        for (Recipe &r : rec)
            for (int &gi : r.gates) gi = take[gi];
        plan.recipes.insert(plan.recipes.end(), rec.begin(), rec.end());
        std::vector<Gate> phys;
        phys.reserve(fused.size());
        for (const Gate &g : fused) phys.push_back(to_physical(g, pi));
        int sperm[MAX_TILE_BITS];
        store_perm(P, S, pi, pi2, n, sperm);
        if (!lower_stage(phys, P, n, o, plan, err, sperm)) return false;
        last_pi = pi;
        last_S = S;
        pi = pi2;
        for (int u = 0; u < n; ++u) inv[pi[u]] = u;
        rem.swap(defer);
    }
    const std::vector<int> end_pi = pi;
    const size_t n_compute = plan.stages.size();
    // restore the identity layout: whole cycles of pi per pass
    for (int guard = 0; guard < 4 * n; ++guard) {
        bool ident = true;
        for (int u = 0; u < n; ++u) ident &= pi[u] == u;
        if (ident) break;
        uint64_t P = low_mask(L);
        std::vector<int> pi2 = pi;
        std::vector<bool> seen(n, false);
        for (int u = 0; u < n; ++u) {
            if (seen[u] || pi[u] == u) continue;
            // cycle through u: logical u sits at pi[u] and belongs at u
            std::vector<int> cyc;
            for (int v = u; !seen[v]; v = inv[v]) {
                seen[v] = true;
                cyc.push_back(v);
            }
            uint64_t cm = 0;
            for (int v : cyc) cm |= (1ull << v) | (1ull << pi[v]);
            if (popc(P | cm) <= m) {
                P |= cm;
                for (int v : cyc) pi2[v] = v;
            } else if (popc(P) + 1 < m) {
                // partial: send the first r logical bits home, the one whose
                // home is taken keeps the cycle going in a shorter form
                std::vector<int> part;
                uint64_t pm = P;
                for (int v : cyc) {
                    uint64_t add = (1ull << v) | (1ull << pi[v]);
                    if (popc(pm | add) > m) break;
                    pm |= add;
                    part.push_back(v);
                }
                if (part.size() < 2) continue;
                // part = v0, v1 = inv[v0], ...; v_i moves home (v_i's home holds
                // v_{i+1}); the last one takes over the physical bit v0 left.
                const int vlast = part.back();
                const int hole = pi[part[0]];
                for (size_t i = 0; i + 1 < part.size(); ++i) pi2[part[i + 1]] = pi[part[i + 1]];
                pi2[part[0]] = part[0];
                for (size_t i = 1; i + 1 < part.size(); ++i) pi2[part[i]] = part[i];
                pi2[vlast] = hole;
                P = pm;
            }
        }
        // bits of P: everything moved must be inside
        uint64_t S = 0;
        for (int u = 0; u < n; ++u)
            if (P >> pi[u] & 1) S |= 1ull << u;
        P = fill_tile(P, m, n);
        S = 0;
        for (int u = 0; u < n; ++u)
            if (P >> pi[u] & 1) S |= 1ull << u;
        int sperm[MAX_TILE_BITS];
        store_perm(P, S, pi, pi2, n, sperm);
        if (!lower_stage({}, P, n, o, plan, err, sperm)) return false;
        pi = pi2;
        for (int u = 0; u < n; ++u) inv[pi[u]] = u;
    }
    for (int u = 0; u < n; ++u)
        if (pi[u] != u) {
            err = "internal: layout restore failed";
            return false;
        }
    if (o.relabel && m < n && n_compute > 0 && plan.stages.size() > n_compute)
        relabel_tail(plan, n, m, L, o, last_pi, last_S, end_pi, n_compute);
    return true;
}

void refresh_plan_mats(Plan &plan, const std::vector<Gate> &gates) {
    for (size_t i = 0; i < plan.ops.size(); ++i) {
        const Recipe &r = plan.recipes[i];
        const DevOp &d = plan.ops[i];
        if (r.gates.size() == 1) {
            const Gate &g = gates[r.gates[0]];
            std::copy(g.m.begin(), g.m.end(), plan.mats.begin() + d.mat);
        } else {
            Block b;
            b.bits = r.bits;
            b.gates = r.gates;
            const Gate g = materialise(b, gates);
            std::copy(g.m.begin(), g.m.end(), plan.mats.begin() + d.mat);
        }
    }
}

void round_mats_f32(const Plan &p, float *out) {
    for (size_t i = 0; i < p.mats.size(); ++i) {
        out[2 * i] = (float)p.mats[i].real();
        out[2 * i + 1] = (float)p.mats[i].imag();
    }
}

namespace {
int lt_slot(int i) { return i ^ ((i >> 5) & 31); }
int lt_spread(int x, int fixed) {
    for (int f = fixed; f; f &= f - 1) {
        const int b = __builtin_ctz(f);
        x = ((x >> b) << (b + 1)) | (x & ((1 << b) - 1));
    }
    return x;
}
}  // namespace

void build_lane_tables(const Plan &plan, const float *hi, const float *lo, std::vector<DevOp> &ops,
                       std::vector<int32_t> &out) {
    constexpr int kLogWaves = 2, EB = 8;  // 4 waves of 64 lanes; complex64 slots
    out.clear();
    ops = plan.ops;
    int32_t next = 0;
    for (DevOp &op : ops) {
        if (op.kind != OPK_MFMA || op.k != 3) continue;
        op.lt = next++;
        const int p0 = op.pos[0], p1 = op.pos[1], p2 = op.pos[2], cl = (int)op.cloc;
        const int fixed = cl | (1 << p0) | (1 << p1) | (1 << p2);
        auto roff = [&](int r) { return ((r & 1) << p0) | ((r >> 1 & 1) << p1) | ((r >> 2 & 1) << p2); };
        auto frag = [&](const float *m, int i16, int g4, int ks) -> float {
            const int kk = 4 * ks + g4;
            const float *u = m + 2 * ((size_t)op.mat + (size_t)(i16 >> 1) * 8 + (kk & 7));
            return kk < 8 ? ((i16 & 1) ? u[1] : u[0]) : ((i16 & 1) ? u[0] : -u[1]);
        };
        for (int t = 0; t < kLaneTableLanes; ++t) {
            const int lane = t & 63, wave = t >> 6, i16 = lane & 15, g4 = lane >> 4;
            const int pl = lt_spread(i16, fixed) | cl;
            int32_t w[kLaneTableWords] = {0};
            w[0] = lt_slot(pl | roff(g4)) * EB;
            w[1] = lt_slot(pl | roff(4 + g4)) * EB;
            w[2] = lt_slot(pl | roff(2 * g4)) * EB;
            w[3] = lt_slot(pl | roff(2 * g4 + 1)) * EB;
            w[4] = lt_slot(lt_spread((wave << 4) | (lane << (4 + kLogWaves)), fixed)) * EB;
            w[5] = lt_slot(lt_spread(lane << (10 + kLogWaves), fixed)) * EB;
            for (int ks = 0; ks < 4; ++ks) {
                const float a = frag(hi, i16, g4, ks), b = lo ? frag(lo, i16, g4, ks) : 0.f;
                std::memcpy(&w[8 + ks], &a, 4);
                std::memcpy(&w[12 + ks], &b, 4);
            }
            out.insert(out.end(), w, w + kLaneTableWords);
        }
    }
}

void round_mats_f32_lo(const Plan &p, const float *hi, float *lo) {
    for (size_t i = 0; i < p.mats.size(); ++i) {
        lo[2 * i] = (float)(p.mats[i].real() - (double)hi[2 * i]);
        lo[2 * i + 1] = (float)(p.mats[i].imag() - (double)hi[2 * i + 1]);
    }
}


// ---------------------------------------------------------------------------
// Register-resident tiles (RegProgram): groups, wave stages, layouts,
// exchanges and fragments
// ---------------------------------------------------------------------------
namespace {

// Positions of a tile bit: lane bits 0..5, then nr register bits, then wave bits.
constexpr int kPosR0 = 6, kMaxPos = 14;
inline bool pos_is_col(int q) { return q < 4; }             // lane bits 0..3: MFMA columns
inline bool pos_is_h(int q) { return q == 4 || q == 5; }   // lane bits 4, 5: MFMA row groups

struct RegCore {
    int k = 0;
    int t[3] = {0, 0, 0};
    bool diag = false;
    int32_t mat = 0;
};

bool reg_core_of(const DevOp &d, RegCore &c) {
    if (d.cloc || d.cglob) return false;
    switch (d.kind) {
        case OPK_MFMA:
        case OPK_DENSEK:
            if (d.k != 3) return false;
            c.diag = false;
            break;
        case OPK_DENSE1:
        case OPK_DENSE2: c.diag = false; break;
        case OPK_DIAG:
            if (d.k < 1 || d.k > 3) return false;
            c.diag = true;
            break;
        default: return false;
    }
    c.k = d.k;
    for (int j = 0; j < d.k; ++j) {
        if (d.pos[j] < 0) return false;
        c.t[j] = d.pos[j];
    }
    c.mat = d.mat;
    return true;
}

// Layout of a tile's bits over positions: at[q] = tile bit at position q,
// where[b] = position of tile bit b.
struct RegLayout {
    int nr = 4;
    int at[kMaxPos];
    int where[kMaxPos];
    bool is_r(int q) const { return q >= kPosR0 && q < kPosR0 + nr; }
    void set(int q, int b) {
        at[q] = b;
        where[b] = q;
    }
    void swap_pos(int q1, int q2) {
        const int b1 = at[q1], b2 = at[q2];
        set(q1, b2);
        set(q2, b1);
    }
    int32_t reg_slot(int j) const {  // LDS byte slot of the register part of v[j]
        int idx = 0;
        for (int r = 0; r < nr; ++r)
            if (j >> r & 1) idx |= 1 << at[kPosR0 + r];
        return lt_slot(idx) * 8;
    }
};

// Cost of bringing target set T to two lane-bit-4/5 positions and one
// register position: 2 per target on a column lane bit (a DPP exchange with
// a register bit), 1 per target moved from a register bit to lane bit 4/5 (a
// permlane swap).
int reg_cost(const RegLayout &L, const int *T) {
    int col = 0, h = 0;
    for (int j = 0; j < 3; ++j) {
        const int q = L.where[T[j]];
        col += pos_is_col(q);
        h += pos_is_h(q);
    }
    return 2 * col + std::max(0, 2 - h);
}

// Emits the exchanges and MFMA micro-ops of cores[b, e) on layout L (and
// their fragment blocks); `avail` = tile bits an op's identity factors may
// use (the layout's lane and register bits).
void schedule_ops(const Plan &plan, const std::vector<RegCore> &cores, int b, int e, const std::vector<int> &avail,
                  RegLayout &L, std::vector<int32_t> &mu, RegProgram &out, uint32_t col_final = 0) {
    // index of the next op using `bit` after op i; bits no later op uses rank
    // last, those the final layout wants on column lane bits (`col_final`)
    // after the others, so they are the first displaced onto column bits
    auto next_use = [&](int bit, int i) {
        for (int j = i + 1; j < e; ++j)
            for (int t = 0; t < cores[j].k; ++t)
                if (cores[j].t[t] == bit) return j;
        return (1 << 30) + (int)(col_final >> bit & 1);
    };
    auto xlane = [&](int qr, int qb) {
        mu.push_back(RG_XLANE | (qr - kPosR0) << 4 | qb << 8);
        L.swap_pos(qr, qb);
        out.exchanges++;
    };
    for (int i = b; i < e; ++i) {
        const RegCore &c = cores[i];
        // targets: the core bits, widened to 3 by the cheapest extra bits
        int T[3] = {c.t[0], c.t[1], c.t[2]};
        if (c.k < 3) {
            int best = 1 << 30;
            int cand[3] = {T[0], T[1], T[2]};
            const int na = (int)avail.size();
            for (int a = 0; a < na; ++a) {
                const int ba = avail[a];
                bool in = false;
                for (int j = 0; j < c.k; ++j) in |= c.t[j] == ba;
                if (in) continue;
                if (c.k == 2) {
                    cand[2] = ba;
                    const int cst = reg_cost(L, cand);
                    if (cst < best) {
                        best = cst;
                        std::copy(cand, cand + 3, T);
                    }
                    continue;
                }
                for (int a2 = a + 1; a2 < na; ++a2) {
                    const int bb = avail[a2];
                    if (bb == c.t[0]) continue;
                    cand[1] = ba;
                    cand[2] = bb;
                    const int cst = reg_cost(L, cand);
                    if (cst < best) {
                        best = cst;
                        std::copy(cand, cand + 3, T);
                    }
                }
            }
        }
        auto is_t = [&](int bit) { return bit == T[0] || bit == T[1] || bit == T[2]; };
        // targets on column lane bits move to register bits (the displaced
        // non-target with the latest next use goes there)
        for (int j = 0; j < 3; ++j) {
            const int q = L.where[T[j]];
            if (!pos_is_col(q)) continue;
            int pick = -1, far = -1;
            for (int r = 0; r < L.nr; ++r) {
                const int bit = L.at[kPosR0 + r];
                if (is_t(bit)) continue;
                const int nu = next_use(bit, i);
                if (nu > far) {
                    far = nu;
                    pick = r;
                }
            }
            xlane(kPosR0 + pick, q);
        }
        // two targets at lane bits 4, 5
        for (;;) {
            int h = 0;
            for (int j = 0; j < 3; ++j) h += pos_is_h(L.where[T[j]]);
            if (h >= 2) break;
            int tr = -1;
            for (int j = 0; j < 3 && tr < 0; ++j)
                if (L.is_r(L.where[T[j]])) tr = j;
            int hq = -1, far = -1;
            for (int q = 4; q <= 5; ++q) {
                if (is_t(L.at[q])) continue;
                const int nu = next_use(L.at[q], i);
                if (nu > far) {
                    far = nu;
                    hq = q;
                }
            }
            xlane(L.where[T[tr]], hq);
        }
        int s = -1;
        for (int j = 0; j < 3; ++j)
            if (L.is_r(L.where[T[j]])) s = L.where[T[j]] - kPosR0;
        mu.push_back(RG_MFMA3 | s << 4);
        out.mfma_ops++;
        // fragments: the 8x8 block over (T[0], T[1], T[2]) as the stacked
        // real 16x16 matrix in this layout's row/column order
        auto tl = [&](int sel, int g) {
            int idx = 0;
            for (int j = 0; j < 3; ++j) {
                const int q = L.where[T[j]];
                const int v = q == 4 ? (g & 1) : q == 5 ? (g >> 1 & 1) : sel;
                idx |= v << j;
            }
            return idx;
        };
        const int kc = c.k, mk = (1 << kc) - 1;
        auto U8 = [&](int o, int in) -> cd {
            if ((o >> kc) != (in >> kc)) return 0.0;
            const int ro = o & mk, ri = in & mk;
            if (c.diag) return ro == ri ? plan.mats[(size_t)c.mat + ro] : cd(0.0);
            return plan.mats[(size_t)c.mat + (size_t)ro * (1 << kc) + ri];
        };
        float R[16][16], Rl[16][16];
        for (int m = 0; m < 16; ++m)
            for (int kk = 0; kk < 16; ++kk) {
                const int go = m >> 2, jj = m & 3, ks = kk >> 2, gi = kk & 3;
                const cd u = U8(tl(jj >> 1, go), tl(ks & 1, gi));
                const bool re_o = !(jj & 1), re_i = ks < 2;
                const double val = re_o ? (re_i ? u.real() : -u.imag()) : (re_i ? u.imag() : u.real());
                R[m][kk] = (float)val;
                Rl[m][kk] = (float)(val - (double)R[m][kk]);
            }
        for (int blk = 0; blk < out.blocks_per_op; ++blk)
            for (int lane = 0; lane < 64; ++lane)
                for (int ks = 0; ks < 4; ++ks)
                    out.frags.push_back(blk ? Rl[lane & 15][4 * ks + (lane >> 4)] : R[lane & 15][4 * ks + (lane >> 4)]);
    }
}

// Wave-tile form of the ops (RG_MFMA3T): a 3-qubit block with all three
// targets on register bits x < y < z, as 16 k-steps of the 4-block
// v_mfma_f32_16x16x1_4b_f32 per group of 8 amplitudes: k-step k reads float
// (k & 1) of the register with bits x, y, z = k >> 1 (the B operand of the
// lane's own column), and accumulator register r returns to the register
// that k = r was read from.  The instruction's layout moves two bits
// between lanes and registers: the outputs of input lane group b (lane bits
// 4, 5) come back in accumulator registers 4b .. 4b + 3, and output rows 4g'
// .. 4g' + 3 land in lane group g'.  So after the op the target o0 sits at
// register bit x, the old occupants of lane bits 4, 5 at register bits y, z,
// and the targets o1, o2 at lane bits 4, 5 (the host orders the matrix rows
// as m = re/im + 2 o0 + 4 o1 + 8 o2 and its columns as k = re/im + 2 t_x +
// 4 t_y + 8 t_z).  A target on a lane bit is first exchanged with a register
// bit (one permlane swap for lane bits 4, 5, a DPP exchange for 0..3).
void schedule_ops_t(const Plan &plan, const std::vector<RegCore> &cores, int b, int e, const std::vector<int> &avail,
                    RegLayout &L, std::vector<int32_t> &mu, RegProgram &out, uint32_t col_final) {
    auto next_use = [&](int bit, int i) {
        for (int j = i + 1; j < e; ++j)
            for (int t = 0; t < cores[j].k; ++t)
                if (cores[j].t[t] == bit) return j;
        return (1 << 30) + (int)(col_final >> bit & 1);
    };
    auto pos_cost = [&](int q) { return L.is_r(q) ? 0 : pos_is_h(q) ? 1 : 3; };
    for (int i = b; i < e; ++i) {
        const RegCore &c = cores[i];
        int T[3] = {c.t[0], c.t[1], c.t[2]};
        if (c.k < 3) {
            // identity factors on the cheapest bits (register-resident first;
            // among those, the bits used soonest: they stay register-resident)
            long best = 1l << 60;
            const int na = (int)avail.size();
            for (int a = 0; a < na; ++a) {
                const int ba = avail[a];
                bool in = false;
                for (int j = 0; j < c.k; ++j) in |= c.t[j] == ba;
                if (in) continue;
                for (int a2 = (c.k == 2 ? a : a + 1); a2 < na; ++a2) {
                    const int bb = avail[a2];
                    if (c.k == 1 && (bb == c.t[0] || bb == ba)) continue;
                    int cand[3] = {c.t[0], c.k == 2 ? c.t[1] : ba, c.k == 2 ? ba : bb};
                    if (c.k == 2 && a2 != a) break;
                    long cst = 0;
                    for (int j = 0; j < 3; ++j) cst += pos_cost(L.where[cand[j]]);
                    cst = cst * 4096 + std::min(next_use(cand[2], i), 4095) + (c.k == 1 ? std::min(next_use(cand[1], i), 4095) : 0);
                    if (cst < best) {
                        best = cst;
                        std::copy(cand, cand + 3, T);
                    }
                }
            }
        }
        auto is_t = [&](int bit) { return bit == T[0] || bit == T[1] || bit == T[2]; };
        // targets on lane bits move to register bits (displacing the
        // non-target used latest)
        for (int j = 0; j < 3; ++j) {
            const int q = L.where[T[j]];
            if (L.is_r(q)) continue;
            int pick = -1, far = -1;
            for (int r = 0; r < L.nr; ++r) {
                const int bit = L.at[kPosR0 + r];
                if (is_t(bit)) continue;
                const int nu = next_use(bit, i);
                if (nu > far) {
                    far = nu;
                    pick = r;
                }
            }
            mu.push_back(RG_XLANE | pick << 4 | q << 8);
            L.swap_pos(kPosR0 + pick, q);
            out.exchanges++;
        }
        int rp[3];
        for (int j = 0; j < 3; ++j) rp[j] = L.where[T[j]] - kPosR0;
        std::sort(rp, rp + 3);
        const int tx = L.at[kPosR0 + rp[0]], ty = L.at[kPosR0 + rp[1]], tz = L.at[kPosR0 + rp[2]];
        // o0 (stays at register bit x): the target used soonest
        int o[3] = {T[0], T[1], T[2]};
        std::sort(o, o + 3, [&](int u, int w) { return next_use(u, i) < next_use(w, i); });
        const int c1 = L.at[4], c2 = L.at[5];
        mu.push_back(RG_MFMA3T | rp[0] << 4 | rp[1] << 8 | rp[2] << 12);
        out.mfma_ops++;
        auto tbit = [&](int bit) {
            for (int j = 0; j < 3; ++j)
                if (T[j] == bit) return j;
            return -1;
        };
        const int jx = tbit(tx), jy = tbit(ty), jz = tbit(tz), j0 = tbit(o[0]), j1 = tbit(o[1]), j2 = tbit(o[2]);
        const int kc = c.k, mk = (1 << kc) - 1;
        auto U8 = [&](int oo, int in) -> cd {
            if ((oo >> kc) != (in >> kc)) return 0.0;
            const int ro = oo & mk, ri = in & mk;
            if (c.diag) return ro == ri ? plan.mats[(size_t)c.mat + ro] : cd(0.0);
            return plan.mats[(size_t)c.mat + (size_t)ro * (1 << kc) + ri];
        };
        float R[16][16], Rl[16][16];
        for (int m = 0; m < 16; ++m)
            for (int k = 0; k < 16; ++k) {
                const int oi = ((m >> 1 & 1) << j0) | ((m >> 2 & 1) << j1) | ((m >> 3 & 1) << j2);
                const int ii = ((k >> 1 & 1) << jx) | ((k >> 2 & 1) << jy) | ((k >> 3 & 1) << jz);
                const cd u = U8(oi, ii);
                const bool re_o = !(m & 1), re_i = !(k & 1);
                const double val = re_o ? (re_i ? u.real() : -u.imag()) : (re_i ? u.imag() : u.real());
                R[m][k] = (float)val;
                Rl[m][k] = (float)(val - (double)R[m][k]);
            }
        for (int blk = 0; blk < out.blocks_per_op; ++blk)
            for (int m = 0; m < 16; ++m)
                for (int k = 0; k < 16; ++k) out.frags.push_back(blk ? Rl[m][k] : R[m][k]);
        // layout after the op
        L.set(kPosR0 + rp[0], o[0]);
        L.set(kPosR0 + rp[1], c1);
        L.set(kPosR0 + rp[2], c2);
        L.set(4, o[1]);
        L.set(5, o[2]);
    }
}

bool stage_cores(const Plan &plan, const DevStage &st, std::vector<RegCore> &cores) {
    cores.clear();
    for (int i = st.op_begin; i < st.op_end; ++i) {
        RegCore c;
        if (!reg_core_of(plan.ops[i], c)) return false;
        cores.push_back(c);
    }
    return true;
}

uint32_t core_mask(const RegCore &c) {
    uint32_t m = 0;
    for (int j = 0; j < c.k; ++j) m |= 1u << c.t[j];
    return m;
}

// Tile bits ordered by first use in cores[b, e) (never used: last, by bit).
std::vector<int> by_first_use(const std::vector<RegCore> &cores, int b, int e, const std::vector<int> &bits) {
    std::vector<int> first(64, 1 << 30);
    for (int i = e - 1; i >= b; --i)
        for (int j = 0; j < cores[i].k; ++j) first[cores[i].t[j]] = i;
    std::vector<int> order = bits;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return first[x] < first[y]; });
    return order;
}

}  // namespace

void build_reg_program(const Plan &plan, bool comp, RegProgram &out) {
    out = RegProgram();
    out.blocks_per_op = comp ? 2 : 1;
    const size_t ns = plan.stages.size();
    out.stage_word.assign(ns, -1);
    out.stage_frag.assign(ns, 0);
    std::vector<RegCore> cores;
    for (size_t si = 0; si < ns; ++si) {
        const DevStage &st = plan.stages[si];
        out.stage_frag[si] = (int32_t)(out.frags.size() / 256);
        if (st.m != kRegTileBits || st.n_grad || st.op_end <= st.op_begin) continue;
        if (!stage_cores(plan, st, cores)) continue;
        const int nops = (int)cores.size();
        // groups: maximal runs of ops whose targets leave two tile bits free (the wave bits)
        std::vector<std::pair<int, int>> groups;
        for (int i = 0; i < nops;) {
            uint32_t U = 0;
            int e = i;
            while (e < nops && popc(U | core_mask(cores[e])) <= kRegTileBits - 2) U |= core_mask(cores[e++]);
            groups.push_back({i, e});
            i = e;
        }
        const size_t w0 = out.words.size();
        out.stage_word[si] = (int32_t)w0;
        out.words.push_back((int32_t)groups.size());
        out.words.push_back(0);  // MFMA ops, filled below
        const int64_t q0 = out.mfma_ops;
        for (const auto &gr : groups) {
            uint32_t U = 0;
            for (int i = gr.first; i < gr.second; ++i) U |= core_mask(cores[i]);
            // wave bits: the two highest bits no op of the group touches
            int wb[2], nw = 0;
            for (int b = kRegTileBits - 1; b >= 0 && nw < 2; --b)
                if (!(U >> b & 1)) wb[nw++] = b;
            std::vector<int> bits;
            for (int b = 0; b < kRegTileBits; ++b)
                if (b != wb[0] && b != wb[1]) bits.push_back(b);
            const std::vector<int> order = by_first_use(cores, gr.first, gr.second, bits);
            // initial layout: earliest-used bits at lane bits 4, 5, then the
            // register bits, the latest-used at the column lane bits
            RegLayout L;
            L.nr = 4;
            const int seq[10] = {4, 5, kPosR0, kPosR0 + 1, kPosR0 + 2, kPosR0 + 3, 0, 1, 2, 3};
            for (int i = 0; i < 10; ++i) L.set(seq[i], order[i]);
            L.set(kPosR0 + 4, wb[0]);
            L.set(kPosR0 + 5, wb[1]);
            const RegLayout L_in = L;
            std::vector<int32_t> mu;
            schedule_ops(plan, cores, gr.first, gr.second, order, L, mu, out);
            out.groups++;
            out.words.push_back((int32_t)mu.size());
            for (int q = 0; q < 6; ++q) out.words.push_back(L_in.at[q]);
            out.words.push_back(L_in.at[kPosR0 + 4]);
            out.words.push_back(L_in.at[kPosR0 + 5]);
            for (int j = 0; j < 16; ++j) out.words.push_back(L_in.reg_slot(j));
            for (int q = 0; q < 6; ++q) out.words.push_back(L.at[q]);
            for (int j = 0; j < 16; ++j) out.words.push_back(L.reg_slot(j));
            out.words.insert(out.words.end(), mu.begin(), mu.end());
        }
        out.words[w0 + 1] = (int32_t)(out.mfma_ops - q0);
    }
}

void build_wave_program(const Plan &plan, bool comp, RegProgram &out) {
    out = RegProgram();
    out.blocks_per_op = comp ? 2 : 1;
    const size_t ns = plan.stages.size(), nall = ns + plan.fb_stages.size();
    out.stage_word.assign(nall, -1);
    out.stage_frag.assign(nall, 0);
    std::vector<RegCore> cores;
    int pending = 0;  // relabelling stages without a program
    for (size_t si = 0; si < nall; ++si) {
        const DevStage &st = si < ns ? plan.stages[si] : plan.fb_stages[si - ns];
        const StageRelabel *rl = si < ns && !plan.relabel.empty() && plan.relabel[si].store_mask
                                     ? &plan.relabel[si] : nullptr;
        pending += rl ? 1 : 0;
        out.stage_frag[si] = (int32_t)(out.frags.size() / 256);
        const int m = st.m;
        if ((m != 11 && m != 12) || st.L < 5 || st.n_grad) continue;
        if (!stage_cores(plan, st, cores)) continue;
        const int nops = (int)cores.size();
        // global bit of tile-local bit b, and of store-order bit j (the j-th
        // lowest bit of the tile, or of the store mask of a relabelling stage)
        auto gbit = [&](int b) { return b < st.L ? b : st.hbits[b - st.L]; };
        int sbits[MAX_TILE_BITS];
        for (int j = 0, c = 0; j < m; ++j) {
            if (!rl) {
                sbits[j] = gbit(j);
                continue;
            }
            while (!(rl->store_mask >> c & 1)) ++c;
            sbits[j] = c++;
        }
        int sperm[MAX_TILE_BITS];
        for (int j = 0; j < m; ++j) sperm[j] = st.sperm[j];
        RegLayout L;
        L.nr = m - 6;
        // load layout (coalesced 16-byte loads, 256 contiguous bytes per 16
        // lanes): tile bit 0 at register bit 0 (the amplitude pair of one
        // load), tile bits 1..4 at lane bits 0..3; the other bits by first
        // use: lane bits 4, 5, then register bits 1..
        L.set(kPosR0, 0);
        for (int q = 0; q < 4; ++q) L.set(q, 1 + q);
        std::vector<int> rest, all;
        for (int b = 0; b < m; ++b) all.push_back(b);
        for (int b = 5; b < m; ++b) rest.push_back(b);
        rest = by_first_use(cores, 0, nops, rest);
        {
            // the first-used bits on register bits 1.. (every op reads its
            // targets from registers), the next two at lane bits 4, 5 (the
            // first op moves them into registers)
            int i = 0;
            for (int r = 1; r < L.nr; ++r) L.set(kPosR0 + r, rest[i++]);
            L.set(4, rest[i++]);
            L.set(5, rest[i++]);
        }
        const RegLayout L_in = L;
        std::vector<int32_t> mu;
        const int64_t q0 = out.mfma_ops;
        const size_t f0 = out.frags.size();
        uint32_t col_final = 0;
        for (int j = 1; j <= 4; ++j) col_final |= 1u << sperm[j];
        schedule_ops_t(plan, cores, 0, nops, all, L, mu, out, col_final);
        if ((out.mfma_ops - q0) * out.blocks_per_op > kWaveMaxFragBlocks) {
            // the stage's fragments would not fit the kernel's LDS: LDS op path
            out.mfma_ops = q0;
            out.frags.resize(f0);
            continue;
        }
        // store layout: store-order bit 0 at register bit 0, store-order bits
        // 1..4 at lane bits 0..3 (any order)
        auto xlane = [&](int qr, int qb) {
            mu.push_back(RG_XLANE | (qr - kPosR0) << 4 | qb << 8);
            L.swap_pos(qr, qb);
            out.exchanges++;
        };
        {
            const int b0 = sperm[0], q = L.where[b0];
            if (L.is_r(q) && q != kPosR0) {
                xlane(q, 4);
                xlane(kPosR0, 4);
            } else if (!L.is_r(q)) {
                xlane(kPosR0, q);
            }
        }
        auto in_set = [&](int b) {
            for (int j = 1; j <= 4; ++j)
                if (sperm[j] == b) return true;
            return false;
        };
        for (int j = 1; j <= 4; ++j) {
            const int b = sperm[j];
            if (pos_is_col(L.where[b])) continue;
            int qc = -1;
            for (int q = 0; q < 4 && qc < 0; ++q)
                if (!in_set(L.at[q])) qc = q;
            if (pos_is_h(L.where[b])) xlane(kPosR0 + 1, L.where[b]);
            xlane(L.where[b], qc);
        }
        const size_t w0 = out.words.size();
        out.stage_word[si] = (int32_t)w0;
        out.words.push_back((int32_t)mu.size());
        out.words.push_back((int32_t)(out.mfma_ops - q0));
        // global bit of each position, load then store layout
        for (int q = 0; q < 6 + L.nr; ++q) out.words.push_back(gbit(L_in.at[q]));
        for (int q = 0; q < 6 + L.nr; ++q) {
            int j = 0;
            while (sperm[j] != L.at[q]) ++j;
            out.words.push_back(sbits[j]);
        }
        out.words.insert(out.words.end(), mu.begin(), mu.end());
        out.words.push_back(rl ? 1 : 0);
        if (rl) {
            out.words.insert(out.words.end(), rl->omap.begin(), rl->omap.end());
            pending--;
        }
        out.groups++;
    }
    out.relabel_ok = pending == 0;
}

}  // namespace desv
