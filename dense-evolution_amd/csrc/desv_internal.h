// Internal structures shared by the host compiler (plan.cpp) and the HIP
// engine (engine.hip).  Not part of the public ABI.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace desv {

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// Host gate representation
// ---------------------------------------------------------------------------
// A gate acts on k target bits (global basis-index bit positions, LSB-first
// matrix index: matrix index bit j <-> bits[j]) and is conditioned on every
// bit of `cmask` being 1.
enum GateKind : int {
    G_DENSE = 0,   // 2^k x 2^k row-major matrix
    G_DIAG = 1,    // 2^k diagonal entries
    G_SCALAR = 2,  // one complex scalar on the whole state (k = 0)
    G_GRAD = 3,    // adjoint inner product Re<lam| M |psi> (M dense or diag)
};

struct Gate {
    int kind = G_DENSE;
    int k = 0;
    int bits[5] = {0, 0, 0, 0, 0};
    uint64_t cmask = 0;
    std::vector<cd> m;   // dense: 4^k entries, diag: 2^k entries, scalar: 1
    bool grad_diag = false;  // for G_GRAD: m holds a diagonal
    int grad_slot = -1;      // for G_GRAD: index into the gradient slots
    int src = -1;            // index of the logical gate it came from

    uint64_t target_mask() const {
        uint64_t t = 0;
        for (int j = 0; j < k; ++j) t |= 1ull << bits[j];
        return t;
    }
    // Bits that must sit inside an LDS tile (non-diagonal action).
    uint64_t tile_mask() const {
        // Gradient probes read their target bits from the tile, diagonal or not.
        if (kind == G_DENSE || kind == G_GRAD) return target_mask();
        return 0;
    }
    uint64_t all_mask() const { return target_mask() | cmask; }
    bool is_diag() const { return kind == G_DIAG || kind == G_SCALAR || (kind == G_GRAD && grad_diag); }
};

// Decode one reference op-table row [g_id, q1, q2, param] into a gate on an
// n-qubit state.  Returns false (with `err`) for unknown ids / bad qubits.
// Identity rows produce no gate (ok = true, out empty).
bool decode_row(const double *row, int n, std::vector<Gate> &out, std::string &err);
// Arbitrary k-qubit matrix (qubits[0] = most significant index bit).
bool decode_matrix(const double *mat, const int *qubits, int k, int n, Gate &g, std::string &err);
// Inverse (conjugate transpose) of a gate.
Gate inverse_gate(const Gate &g);
// Generator gate for the adjoint sweep of a theta-carrying row:
// grad contribution = Re<lam| M |psi> with M = 2 * dU/dtheta * U^dag.
bool grad_gate_for_row(const double *row, int n, int slot, Gate &g, std::string &err);

// ---------------------------------------------------------------------------
// Device-side op encoding (one LDS-tile micro-op)
// ---------------------------------------------------------------------------
enum OpKind : int32_t {
    OPK_DENSE1 = 0,   // 2x2 on one tile bit
    OPK_DENSE2 = 1,   // 4x4 on two tile bits
    OPK_DENSEK = 2,   // 2^k x 2^k (k = 3..5) on VALU
    OPK_DIAG = 3,     // diagonal over k bits (tile-local or global)
    OPK_SCALAR = 4,   // whole-tile scalar
    OPK_GRAD = 5,     // adjoint inner product (dense k<=2 on tile bits)
    OPK_GRADDIAG = 6, // adjoint inner product, diagonal generator
    OPK_MFMA = 7,     // 2^k x 2^k (k = 3..5) complex GEMM on MFMA (complex64)
};

struct DevOp {
    int32_t kind;
    int32_t k;
    int32_t mat;       // offset (complex elements) into the matrix pool
    int32_t slot;      // gradient slot (program-wide index)
    int32_t pos[5];    // tile-local bit of matrix index bit j; for OPK_DIAG a
                       // negative value -1-b means global bit b (outside tile)
    uint32_t cloc;     // controls on tile-local bits
    int32_t lt = -1;   // 3-qubit MFMA ops: index of the op's per-lane table
                       // (build_lane_tables), -1: computed in the kernel
    uint64_t cglob;    // controls on global bits outside the tile
};

// Per-lane constants of a 3-qubit MFMA op for a 256-lane workgroup, four
// 16-byte words per lane: LDS byte slots (B loads rows g, 4+g; C stores rows
// 2g, 2g+1), the column-block slot tables of the lane (block g and 64 g), and
// the stacked real A fragments (hi parts, then lo parts of compensated
// matrices).  Computed once per program on the host instead of in every
// workgroup of every stage.
constexpr int kLaneTableWords = 16;  // int32 per lane
constexpr int kLaneTableLanes = 256;

constexpr int MAX_HBITS = 24;
constexpr int MAX_TILE_BITS = 16;

// Register groups: a complex64 forward stage of 2^12-amplitude tiles whose
// ops are all uncontrolled dense or diagonal blocks of at most 3 bits can run
// with the tile in registers instead of LDS: each of the 4 waves holds 1024
// amplitudes (64 lanes x 16 registers), so a tile-local bit sits at one of 12
// positions -- lane bits 0..5, register bits 0..3 (amplitude v[j] of a lane,
// bit r of j) or wave bits 0..1.  A group reads the tile from LDS in a layout
// (which bit sits where), runs its ops on the registers and writes the tile
// back in the layout it ends in; no op of a group touches its wave bits, so
// a group needs no barrier until its write-back.  Every op of a group is a
// 3-bit block on MFMA (blocks of 1-2 bits are widened by identity factors on
// other bits) whose targets sit at lane bits 4, 5 and one register bit; the
// host inserts lane/register bit exchanges in front of each op to get them
// there.  Stage words (int32):
//   [0] groups, [1] MFMA ops, then per group:
//   [0] micro-ops, [1..6] tile bit at lane bits 0..5 (read layout),
//   [7..8] tile bit at wave bits 0..1, [9..24] byte slot of register j (read
//   layout), [25..30] tile bit at lane bits 0..5 (write layout), [31..46]
//   byte slot of register j (write layout), then one word per micro-op:
//   RG_XLANE | r << 4 | b << 8   exchange register bit r with lane bit b
//   RG_MFMA3 | s << 4            3-bit block: targets at lane bits 4, 5 and
//                                register bit s; A fragments from the next
//                                fragment block of the stage
// Fragment blocks: 64 lanes x 4 floats (the stacked real A fragments of
// ks = 0..3), hi parts, then (compensated programs) lo parts, per MFMA op.
constexpr int kRegTileBits = 12;
constexpr int kRegGroupHeader = 47;
enum RegCode : int32_t { RG_XLANE = 1, RG_MFMA3 = 2, RG_MFMA3T = 3 };
// Wave-tile stages keep their fragment blocks in LDS (one copy per workgroup
// of 4 waves): at most this many 1 KiB blocks per stage.
constexpr int kWaveMaxFragBlocks = 64;
struct RegProgram {
    std::vector<int32_t> words;       // all register stages' words
    std::vector<int32_t> stage_word;  // per stage: offset into words, -1 = LDS op path
    std::vector<int32_t> stage_frag;  // per stage: index of its first fragment block
    std::vector<float> frags;         // fragment blocks, 256 floats each
    int blocks_per_op = 1;            // 2 for compensated programs (hi, lo)
    int64_t groups = 0, exchanges = 0, mfma_ops = 0;
    // wave programs only: entries past plan.stages.size() belong to
    // plan.fb_stages; relabel_ok = every relabelling stage has a program
    bool relabel_ok = true;
};

struct DevStage {
    int32_t m;         // tile bits (amplitudes per tile = 2^m)
    int32_t L;         // low contiguous global bits inside the tile
    int32_t nh;        // high tile bits (m - L)
    int32_t hbits[MAX_HBITS];
    uint64_t tile_mask;
    int64_t n_tiles;
    int32_t op_begin, op_end;
    int32_t n_grad;    // OPK_GRAD* ops in this stage
    int32_t heavy;     // kernel variant: 0 light, 1 k >= 3 VALU blocks, 2 MFMA blocks (k <= 4), 3 MFMA up to k = 5
    // Write-back bit permutation: store-order bit j (the j-th lowest global
    // bit of the tile) takes tile-local bit sperm[j].  Identity unless
    // `permuted`; lets a stage move qubits between the low (contiguous) index
    // bits and high bits at no extra HBM cost.
    int32_t permuted;
    int32_t sperm[MAX_TILE_BITS];
};

// How a device op's matrix is formed from the input gate list: a lone gate
// (gates.size() == 1, bits == 0) or the product of `gates` (in order) over the
// logical bit set `bits`.  Lets a compiled plan refresh its matrices for new
// angles without rescheduling.
struct Recipe {
    std::vector<int> gates;
    uint64_t bits = 0;
};

// A stage that writes out of place (into the second state buffer) with its
// outer bits relabelled: store-order bit j of the tile goes to the j-th
// lowest bit of `store_mask` (taking tile-local bit sperm[j] as usual), and
// the i-th lowest bit outside the tile (load side) goes to bit omap[i].  One
// such stage ends a program in the identity layout with no restore passes:
// only the new low bits need to be inside its tile, the other qubits move
// with the tile addresses.
struct StageRelabel {
    uint64_t store_mask = 0;  // 0: an ordinary in-place stage
    std::vector<int> omap;
};

struct Plan {
    int n = 0;
    std::vector<Recipe> recipes;  // one per op (empty for restore-only stages)
    int64_t n_gates = 0;
    int64_t mfma_blocks = 0;
    int n_grad_slots = 0;
    std::vector<DevStage> stages;
    std::vector<DevOp> ops;
    std::vector<cd> mats;  // pooled matrices (complex128), rounded at upload
    // Out-of-place tail (CompileOpts::relabel): `relabel` has one entry per
    // stage when some stage relabels; such a plan also keeps the in-place
    // form of its tail, `fb_stages`, which replaces stages[fb_from ..] when no
    // second buffer is available.  fb_from = -1: no relabelled stages.
    std::vector<StageRelabel> relabel;
    int fb_from = -1;
    std::vector<DevStage> fb_stages;
};

struct CompileOpts {
    int fuse_bits = 2;
    int tile_bits = 12;
    int low_bits = 5;
    int per_gate = 0;
    int use_mfma = 0;
    int compensated = 0;   // complex64: matrices as FP32 hi + lo pairs, lo applied first
    int nv = 1;            // vectors per tile (2 for the adjoint reverse sweep)
    int remap = 1;         // let stages permute which qubits sit on the low bits
    int window = 4096;     // gates scanned ahead when choosing a tile
    int reg_groups = 0;    // complex64 MFMA programs: 1 register-group stages, 2 wave-tile stages (RegProgram)
    int relabel = 0;       // end with an out-of-place relabelling stage instead of restore passes (Plan::relabel)
    int search = 0;        // randomised planner restarts; the cheapest plan is kept
};

// Build the device plan for a gate list (in application order).
bool build_plan(const std::vector<Gate> &gates, int n, const CompileOpts &o, Plan &plan,
                std::string &err);
// Record the thread-local message returned by desv_last_error(); returns code.
int set_last_error(int code, const std::string &msg);

// Recompute plan.mats from a gate list with the same structure (same kinds,
// bits and controls; only matrix entries differ) as the one compiled.
void refresh_plan_mats(Plan &plan, const std::vector<Gate> &gates);

// plan.mats rounded to complex64 (interleaved floats, 2 per entry), round to
// nearest (products were formed in double).
void round_mats_f32(const Plan &p, float *out);
// lo[i] = the FP32 rounding of plan.mats[i] - hi[i] (hi from round_mats_f32):
// hi + lo carries each entry to about 2^-48 relative.
void round_mats_f32_lo(const Plan &p, const float *hi, float *lo);

// Fill the per-lane tables of every 3-qubit MFMA op of `plan` (complex64
// matrices hi, lo as uploaded; lo may be null) into `out` and set the ops'
// `lt` indices in `ops` (a copy of plan.ops).
void build_lane_tables(const Plan &plan, const float *hi, const float *lo, std::vector<DevOp> &ops,
                       std::vector<int32_t> &out);

// Register-group programs of every eligible stage of `plan` (see RegProgram;
// fragments rounded from plan.mats, with lo parts when `comp`).
void build_reg_program(const Plan &plan, bool comp, RegProgram &out);
// Wave-tile programs: every eligible stage of 2^11- or 2^12-amplitude tiles
// runs one tile per wave, loaded from HBM into registers and stored back
// (the stage's A fragments in LDS).  Stage words: [0] micro-ops, [1] MFMA
// ops, [2 .. 2+m) global bit of each position (lane bits 0..5, register bits
// 0..m-7) in the load layout, [2+m .. 2+2m) the same for the store layout,
// then the micro-ops: RG_XLANE as for groups, RG_MFMA3T | x << 4 | y << 8 |
// z << 12 a 3-qubit block on register bits x < y < z in the 4-block MFMA
// form (plan.cpp, schedule_ops_t; fragment block = 16 rows x 16 k-steps,
// row-major); then 0 for an in-place stage, or 1 and the n - m destination
// bits of the outer bits (StageRelabel::omap) for a relabelling one.  The
// load layout holds tile bit 0 at
// register bit 0 and tile bits 1..4 at lane bits 0..3; the store layout
// store-order bit 0 at register bit 0 and store-order bits 1..4 at lane bits
// 0..3 (16-byte accesses, 256 contiguous bytes per 16 lanes).  Programs are
// built for plan.stages, then plan.fb_stages.
void build_wave_program(const Plan &plan, bool comp, RegProgram &out);

}  // namespace desv
