// cgppd_hip.hip -- gfx950 kernels and C ABI for cgppd's REMC energy hot path.
//
// Design (one GPU, HBM-resident replicas):
//   * Topology (molecule offsets, residue types) is shared by all replicas;
//     per replica the device holds float4 positions {x,y,z,word}, float4
//     relative positions, float4 molecule centres, the molecule-pair energy
//     matrix (FP64, kcal/mol), the running total, temperature, step sizes,
//     move counters and the Philox step counter.  Within each molecule the
//     device order puts charged residues first, then sorts by type, so that
//     a molecule is a few constant-type runs.
//   * Full energies (full_energy_kernel, intra_energy_kernel) and the MC
//     delta-E rows share the FP32 pair terms with per-lane Kahan sums and
//     FP64 reductions per molecule pair.
//   * An MC step of every replica: the row kernel (mc_row_kernel, j
//     residues in the lanes, the moved molecule's residues read as
//     wave-uniform records; or mc_pair_kernel, i residues in the lanes, for
//     systems with flexible molecules) evaluates each replica's pending
//     proposal against all other molecules into fixed-point row buffers;
//     the finish kernel (mc_finish_rigid_kernel / mc_finish_kernel) takes
//     the Metropolis decision in FP64, applies the move and lays out the
//     next proposal.  Rigid systems run speculative windows of up to 8
//     steps per launch pair instead (mc_row_kernel<., ., true> and
//     mc_window_kernel); small rigid systems run many steps per launch in
//     one workgroup per replica (mc_resident_kernel).
//   * Exchange intervals are captured once in a hipGraph and replayed; the
//     autotuner picks the layout, grid and window size.
//   * No MFMA in the pair sums (a table-lookup n-body reduction); the
//     learned surrogate's MLP kernels (mlp.h) run on BF16 MFMA.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cgppd_hip.h"
#include "topology.h"
#include "mlp.h"

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerWG = 4;
constexpr int kWG = kWave * kWavesPerWG;  // threads per workgroup
constexpr int kJChunk = kWG;              // j residues staged per item
// j residues per intra-molecular item of a flexible molecule (32 per wave):
// half the i-lane chunk, so that a proposal's intra items spread over twice
// as many workgroups
constexpr int kIntraChunk = kWG / 2;
__host__ __device__ constexpr int intra_chunks(int n) { return (n + kIntraChunk - 1) / kIntraChunk; }
constexpr int kIBlock = kWave;            // i residues per item (one per lane)
constexpr int kRowShift = 9;              // LDS table row stride: 32 entries x 16 B
constexpr int kColShift = 4;              // 16-byte entries
constexpr int kTableBytes = CGPPD_AA_COUNT << kRowShift;
constexpr uint32_t kRowMask = 0x1Fu << kRowShift;
constexpr uint32_t kChargedBit = 0x80000000u;
// Residue word (the w lane of positions): type << kRowShift | charged bit |
// (residues left in the residue's constant-type run, itself included,
// capped at kRunMax) << kRunShift.
constexpr int kRunShift = 14;
constexpr uint32_t kRunMax = 0x1FFFFu;
// Fixed-point row partials: 2^-32 kcal/mol units in 64-bit integers, so
// that their atomic sum does not depend on the order of the additions.
// Partials of magnitude 2^17 kcal/mol or more (clashes) take the FP64 path.
constexpr double kFixScale = 4294967296.0;
constexpr double kFixInv = 1.0 / 4294967296.0;
constexpr double kFixMax = 131072.0;
constexpr uint32_t kTagMC = 0x4D430000u;
// How the pair kernel evaluates a flexible molecule's internal energy for a
// pending proposal (PairParams::prop_moved.x).  kMovedOne: a local
// translate or crankshaft whose centre shift did not wrap the molecule, so
// every other residue keeps its stored coordinates bit for bit; only the
// moved residue's pairs and the bonded terms are evaluated, new minus
// current.  kMovedAll: every other proposal rewrites the coordinates of many
// residues (rounded), and the whole internal energy is recomputed, as the
// reference does.
constexpr int kMovedNone = 0;  // a molecule without internal energy
constexpr int kMovedOne = 2;   // molecule-local device index in .y
constexpr int kMovedAll = 3;
constexpr uint32_t kTagEX = 0x45580000u;
constexpr uint32_t kTagSG = 0x53470000u;  // surrogate minibatch draws
constexpr uint32_t kTagDA = 0x44410000u;  // delayed-acceptance stage-one uniforms
#define CGPPD_MAX_MOLECULES 1024  // molecules per system
#define CGPPD_MAX_REPLICAS 1024   // replicas per context (LDS item offsets of the pair kernel)

// Model constants, as the reference evaluates them in float
// (src/constants.h:8-26): T_room*Rgas/kcal and e^2 N_A/(4 pi eps0 kcal Ang D).
constexpr float kRgas = 8.3144621f;
constexpr float kKcal = 4184.0f;
constexpr float kLambda = 0.159f;
constexpr float kE0 = -2.27f;
// exp(-r/xi) = exp2(r * kDHExp2), xi = 10 A
constexpr float kDHExp2 = -0.14426950408889634f;

struct Coef {
    float sig2;  // sigma_ij^2
    float A;     // -4 E_ij (raw RT) or -4 E_ij * RT_to_kcal (scaled)
    float rc2;   // (sigma_ij * 2^(1/6))^2 when E_ij > 0, else -1
    float qq;    // q_i q_j (raw) or q_i q_j * DH factor (scaled)
};

// Pair coefficients of the j-lane row kernel (kcal/mol), in the form that
// needs the fewest operations per pair: with u = 1/r^6,
//   attractive branch  A (s^12 - s^6) = u (B u - C),  B = A sigma^12, C = A sigma^6
//   repulsive branch   -A/2 - A (s^12 - s^6) = K - u (B u - C),  K = -A/2
// (A = -4 E_ij RT_to_kcal).  The repulsive branch holds where E_ij > 0 and
// r < 2^(1/6) sigma, i.e. (sigma/r)^6 > 1/2; with A < 0 that is
// C u + K < 0.  Pairs with E_ij <= 0 store K = FLT_MAX, for which the test
// never holds (K itself is then unused).
struct RowCoef {
    float B, C;
    float qq;    // q_i q_j DH factor
    float K;
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11)
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

__host__ __device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint32_t lo0 = 0xD2511F53u * c0;
        uint32_t hi0 = mulhi32(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2;
        uint32_t hi1 = mulhi32(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

__device__ __forceinline__ double unit32(uint32_t w) { return (double)w * (1.0 / 4294967296.0); }
__device__ __forceinline__ double unit53(uint32_t hi, uint32_t lo)
{
    return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

struct Proposal {
    int mol;
    int rotate;
    uint32_t kind_word;  // picks translate/rotate/flex/local for flexible molecules
    double axis[3];  // raw (u-0.5, 0.5-u, u-0.5)
    double u;        // Metropolis uniform
};

// The proposal of MC step `step` of replica r.  Molecule pick as
// src/Replica.cpp:355, Bernoulli(0.5) translate/rotate as
// src/Molecule.cpp:461, axis draw as src/Molecule.cpp:210-221.
__device__ __forceinline__ Proposal draw_proposal(uint64_t seed, int r, int64_t step, int n_mol)
{
    uint32_t w[4], v[4];
    const uint32_t s_lo = (uint32_t)step, s_hi = (uint32_t)((uint64_t)step >> 32);
    philox(s_lo, s_hi, (uint32_t)r, kTagMC, (uint32_t)seed, (uint32_t)(seed >> 32), w);
    philox(s_lo, s_hi, (uint32_t)r, kTagMC | 1u, (uint32_t)seed, (uint32_t)(seed >> 32), v);
    Proposal p;
    p.mol = (int)(((uint64_t)w[0] * (uint64_t)n_mol) >> 32);
    p.rotate = w[1] < 0x80000000u ? 1 : 0;
    p.kind_word = w[1];
    p.axis[0] = unit32(v[0]) - 0.5;
    p.axis[1] = 0.5 - unit32(v[1]);
    p.axis[2] = unit32(v[2]) - 0.5;
    p.u = unit53(w[2], w[3]);
    return p;
}

// The molecule draw_proposal picks at MC step `step` of replica r (the
// host replays it to count the windows of a run).
__host__ __device__ __forceinline__ int draw_mol(uint64_t seed, int r, int64_t step, int n_mol)
{
    uint32_t w[4];
    philox((uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)r, kTagMC, (uint32_t)seed,
           (uint32_t)(seed >> 32), w);
    return (int)(((uint64_t)w[0] * (uint64_t)n_mol) >> 32);
}

// Speculative windows of MC steps (rigid systems).  A window is a run of
// consecutive MC steps of one replica that move distinct molecules: each
// step's proposal then depends only on the state at the window's start
// (its molecule has not moved earlier in the window), so the delta-E rows
// of all of them are evaluated in one launch against that state, together
// with the energies between the proposals themselves, and the decisions
// are then taken one after another in step order (a proposal's row entry
// for a molecule accepted earlier in the window is replaced by the
// proposal-proposal energy).  The decisions, coordinates and energies are
// those of the one-step-per-launch chain.  A window holds at most kMaxWin
// steps, ends before the first repeated molecule and never crosses `stop`
// (the end of the caller's run of steps).  mol_at(k) is the molecule of
// step `step + k`.
constexpr int kMaxWin = 8;
template <typename F>
__host__ __device__ __forceinline__ int window_len(F mol_at, int64_t step, int64_t stop, int W)
{
    int m[kMaxWin];
    int len = 0;
#pragma unroll
    for (int k = 0; k < kMaxWin; ++k) {
        if (k >= W || step + k >= stop) break;
        const int x = mol_at(k);
        bool seen = false;
#pragma unroll
        for (int q = 0; q < kMaxWin; ++q) seen |= q < k && m[q] == x;
        if (seen) break;
        m[k] = x;
        len = k + 1;
    }
    return len;
}

struct Move {
    int rotate;
    float c[3];   // new centre
    double q[4];  // normalised quaternion (w, x, y, z) for rotations
};

// Builds the rigid move from a proposal.  Translation: src/Molecule.cpp:262-266
// (float unit vector scaled by the step, src/vector3f.cpp:58-62,75-78), new
// centre wrapped with fmod (src/Molecule.cpp:146-159,234-241).  Rotation:
// src/Molecule.cpp:193-208,216-221 with src/Quaternion.cpp:12-38.  No FMA
// contraction, so the result is bit-identical to the CPU restatement.
__device__ Move make_move(const Proposal &p, float4 c_old, float tstep, double rcos, double rsin, float box)
{
#pragma clang fp contract(off)
    Move mv;
    mv.rotate = p.rotate;
    if (!p.rotate) {
        float n0 = (float)p.axis[0], n1 = (float)p.axis[1], n2 = (float)p.axis[2];
        float mag = sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
        if (mag != 0.0f) {
            n0 = n0 / mag;
            n1 = n1 / mag;
            n2 = n2 / mag;
        }
        mv.c[0] = fmodf(c_old.x + (float)((double)n0 * (double)tstep), box);
        mv.c[1] = fmodf(c_old.y + (float)((double)n1 * (double)tstep), box);
        mv.c[2] = fmodf(c_old.z + (float)((double)n2 * (double)tstep), box);
        mv.q[0] = 1.0;
        mv.q[1] = mv.q[2] = mv.q[3] = 0.0;
    } else {
        double ax = p.axis[0], ay = p.axis[1], az = p.axis[2];
        double m = sqrt(ax * ax + ay * ay + az * az);
        ax /= m;
        ay /= m;
        az /= m;
        double w = rcos, x = rsin * ax, y = rsin * ay, z = rsin * az;
        double qm = sqrt(w * w + x * x + y * y + z * z);
        mv.q[0] = w / qm;
        mv.q[1] = x / qm;
        mv.q[2] = y / qm;
        mv.q[3] = z / qm;
        mv.c[0] = c_old.x;
        mv.c[1] = c_old.y;
        mv.c[2] = c_old.z;
    }
    return mv;
}

// New relative and absolute position of one residue under the move:
// src/Quaternion.cpp:95-112 and src/Residue.cpp:93-99 for rotations,
// src/Molecule.cpp:183-191 for translations.
__device__ void move_residue(const Move &mv, float4 rel, float4 &rel_new, float4 &pos_new)
{
#pragma clang fp contract(off)
    if (!mv.rotate) {
        rel_new = rel;
        pos_new.x = rel.x + mv.c[0];
        pos_new.y = rel.y + mv.c[1];
        pos_new.z = rel.z + mv.c[2];
    } else {
        const double w = mv.q[0], x = mv.q[1], y = mv.q[2], z = mv.q[3];
        const double w2 = w * w, x2 = x * x, y2 = y * y, z2 = z * z;
        const double wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
        const double vx = rel.x, vy = rel.y, vz = rel.z;
        rel_new.x = (float)(vx * (w2 + x2 - y2 - z2) + (vy * (xy - wz) + vz * (wy + xz)) * 2.0);
        rel_new.y = (float)((vx * (wz + xy) + vz * (yz - wx)) * 2.0 + vy * (w2 - x2 + y2 - z2));
        rel_new.z = (float)((vx * (xz - wy) + vy * (yz + wx)) * 2.0 + vz * (w2 - x2 - y2 + z2));
        rel_new.w = rel.w;
        pos_new.x = mv.c[0] + rel_new.x;
        pos_new.y = mv.c[1] + rel_new.y;
        pos_new.z = mv.c[2] + rel_new.z;
    }
    pos_new.w = rel.w;
}

// The Philox draw of one MC step with its molecule's residue range (the
// window kernel lays these out one window ahead).
struct WinDraw {
    int mol, rotate, beg, n;  // molecule, rotation (else translation), first residue, residues
    float t[3];               // translation: the float increments of the centre
    double q[4];              // rotation: the normalised quaternion
    double u;                 // Metropolis uniform
};

// The parts of make_move that do not depend on the molecule's centre
// (the same operations, so that window moves are bit-identical to it).
__device__ void move_direction(const Proposal &p, float tstep, double rcos, double rsin, WinDraw &d)
{
#pragma clang fp contract(off)
    d.rotate = p.rotate;
    d.t[0] = d.t[1] = d.t[2] = 0.0f;
    d.q[0] = 1.0;
    d.q[1] = d.q[2] = d.q[3] = 0.0;
    if (!p.rotate) {
        float n0 = (float)p.axis[0], n1 = (float)p.axis[1], n2 = (float)p.axis[2];
        float mag = sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
        if (mag != 0.0f) {
            n0 = n0 / mag;
            n1 = n1 / mag;
            n2 = n2 / mag;
        }
        d.t[0] = (float)((double)n0 * (double)tstep);
        d.t[1] = (float)((double)n1 * (double)tstep);
        d.t[2] = (float)((double)n2 * (double)tstep);
    } else {
        double ax = p.axis[0], ay = p.axis[1], az = p.axis[2];
        double m = sqrt(ax * ax + ay * ay + az * az);
        ax /= m;
        ay /= m;
        az /= m;
        double w = rcos, x = rsin * ax, y = rsin * ay, z = rsin * az;
        double qm = sqrt(w * w + x * x + y * y + z * z);
        d.q[0] = w / qm;
        d.q[1] = x / qm;
        d.q[2] = y / qm;
        d.q[3] = z / qm;
    }
}

// The move of a window draw about the molecule's centre c_old (make_move).
__device__ Move move_at(const WinDraw &d, float4 c_old, float box)
{
#pragma clang fp contract(off)
    Move mv;
    mv.rotate = d.rotate;
    if (!d.rotate) {
        mv.c[0] = fmodf(c_old.x + d.t[0], box);
        mv.c[1] = fmodf(c_old.y + d.t[1], box);
        mv.c[2] = fmodf(c_old.z + d.t[2], box);
    } else {
        mv.c[0] = c_old.x;
        mv.c[1] = c_old.y;
        mv.c[2] = c_old.z;
    }
    mv.q[0] = d.q[0];
    mv.q[1] = d.q[1];
    mv.q[2] = d.q[2];
    mv.q[3] = d.q[3];
    return mv;
}

// The quaternion's rotation terms as move_residue forms them, once per move.
struct RotTerms {
    double a, b, c, d, e, f, g, h, i;
};
__device__ RotTerms rot_terms(const Move &mv)
{
#pragma clang fp contract(off)
    const double w = mv.q[0], x = mv.q[1], y = mv.q[2], z = mv.q[3];
    const double w2 = w * w, x2 = x * x, y2 = y * y, z2 = z * z;
    const double wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
    RotTerms k;
    k.a = w2 + x2 - y2 - z2;
    k.b = xy - wz;
    k.c = wy + xz;
    k.d = wz + xy;
    k.e = yz - wx;
    k.f = w2 - x2 + y2 - z2;
    k.g = xz - wy;
    k.h = yz + wx;
    k.i = w2 - x2 - y2 + z2;
    return k;
}

// move_residue with the rotation terms formed beforehand (the same
// operations on the same values).
__device__ __forceinline__ void move_residue_terms(const Move &mv, const RotTerms &k, float4 rel, float4 &rel_new,
                                                   float4 &pos_new)
{
#pragma clang fp contract(off)
    if (!mv.rotate) {
        rel_new = rel;
        pos_new.x = rel.x + mv.c[0];
        pos_new.y = rel.y + mv.c[1];
        pos_new.z = rel.z + mv.c[2];
    } else {
        const double vx = rel.x, vy = rel.y, vz = rel.z;
        rel_new.x = (float)(vx * k.a + (vy * k.b + vz * k.c) * 2.0);
        rel_new.y = (float)((vx * k.d + vz * k.e) * 2.0 + vy * k.f);
        rel_new.z = (float)((vx * k.g + vy * k.h) * 2.0 + vz * k.i);
        rel_new.w = rel.w;
        pos_new.x = mv.c[0] + rel_new.x;
        pos_new.y = mv.c[1] + rel_new.y;
        pos_new.z = mv.c[2] + rel_new.z;
    }
    pos_new.w = rel.w;
}

// ---------------------------------------------------------------------------
// Kernel parameters
// ---------------------------------------------------------------------------
struct PairParams {
    // topology
    int n_mol, n_res, stride;
    int n_active;       // molecules taking part (n_mol, or n_noncrowder)
    int crowd_mol;      // repulsive crowding: first crowder molecule (n_mol when off)
    int ib_max;        // i-blocks of the largest molecule
    int jc_count;       // j-chunks over n_res
    const int *mol_off;  // n_mol + 1
    const int *mol_of;   // n_res
    const int *run_of;       // n_res: run (constant type within a molecule) of each residue
    const int *run_start;    // n_runs + 1: first residue of each run
    const unsigned *run_code;  // n_runs: the run's residue code (table row | charged flag)
    // replicas
    int rep_begin, rep_count;
    // MC kernels of one replica group see its replicas as 0..rep_count-1
    // (the per-replica pointers start at the group's first replica);
    // rep_rng is that replica's index in the context, which keys the
    // Philox streams so that the grouping leaves every trajectory unchanged
    int rep_rng;
    int xcd_map;  // items of replica r stay on XCD r % 8
    const Coef *table;   // 20 x 20 entries, row = j type
    const Coef *table_scaled;  // the same in kcal/mol (finish kernels' bonded terms)
    float box, inv_box;
    float4 *pos;         // R * stride
    float4 *rel;         // R * stride
    float4 *center;      // R * n_mol
    float *radius;       // R * n_mol, max |relative position| of each molecule
    int fast_image;      // use the per-segment image shift where it is exact
    float lj_reach;      // above every pair type's switch distance (2^(1/6) sigma_ij of E_ij > 0)
    float half_box;      // box / 2 and 0.45 box (kernel arguments, so that the row kernel keeps
    float image_reach;   // them in scalar registers)
    // full-energy outputs (raw LJ/DH)
    double *out_lj, *out_dh;  // R
    double *out_mat;          // R * M * M (may be null)
    float rt_to_kcal, dh_factor;
    const RowCoef *row_table;  // 20 x 20, entry tj * 20 + ti
    float charge[CGPPD_AA_COUNT];
    // MC state
    uint64_t seed;
    double *rowbuf;       // R * M
    unsigned long long *rowfix;  // R * M: row partials in fixed point (kFixScale units, two's complement)
    unsigned *done;       // R
    double *emat;         // R * M * M
    double *energy;       // R
    const double *temperature;
    const float *tstep;
    const double *rcos, *rsin;
    long long *counters;  // R * 3
    long long *steps;     // R
    unsigned long long *pair_evals;  // 1: inter-molecular residue pairs evaluated by MC steps
    // the pending proposal of every replica (laid out by mc_finish_kernel)
    int max_n;            // residues of the largest molecule (prop_* row stride)
    float4 *prop_pos;     // R * max_n: proposed positions, device order, w = code
    float4 *prop_rel;     // R * max_n: proposed relative positions
    float4 *prop_center;  // R
    float *prop_radius;   // R: bounding radius of the proposed molecule
    int4 *prop_moved;     // R: {kind, moved residue, 0, 0} (flexible molecules)
    int *prop_mol;        // R
    int *prop_nib;        // R: i-blocks of the moved molecule
    int *prop_n;          // R: residues of the moved molecule
    int *prop_items;      // R: items of the i-lane pair kernel for the pending proposal
    int4 *prop_hdr;       // R: {molecule, first residue, residues, bounding radius bits} of the proposal
    int row_j;            // j residues per j-block of the row kernel (64 x residues per lane)
    // speculative windows (win_log2 = 0: one MC step per launch).  Slot s
    // of replica r is window entry v = r W + s (W = 1 << win_log2); slot 0's
    // proposal is the pending proposal above (prop_pos, prop_hdr, ...), the
    // others live in wprop_pos / wprop_rel.
    int win_log2;
    int4 *whdr;           // R W: {molecule, first residue, residues, bounding radius bits}
    float4 *wcen;         // R W: proposed centre; w = residues of the earlier slots (int bits)
    int *wlen;            // R: slots of the current window (0: the replica has reached wstop)
    long long *wstop;     // R: step at which the current run of steps ends
    float4 *wprop_pos;    // R (W - 1) max_n (+ 8 records of padding): slots 1..W-1
    float4 *wprop_rel;    // R (W - 1) max_n
    // window rows: entry v (M + W) + J, J < M the molecule row, M + q the
    // energy between slot v's and slot q's proposals (q < s)
    unsigned long long *wrowfix;
    double *wrowbuf;
    double *wu;                  // R W: Metropolis uniforms of the current window's steps
    struct WinDraw *wd;          // R W: draws of the W steps after the current window
    // flexible molecules (null / 0 when the system has none).  Residue
    // indices are device indices (global); lists are CSR per molecule.
    int n_flex;                 // number of flexible molecules
    // 1: flexible molecules carry their internal energy (intra-molecular
    // pairs and bonded terms); 0 under the rigid-only override
    int internal;
    const int *flex_mols;       // n_flex: the flexible molecules
    const int *flex_item_off;   // n_flex + 1: full-energy items per replica
    const int *mol_flex;        // M: 1 = flexible
    const int4 *rtag;           // N: chain key, rigid-domain id, segment-bond id, caller index in the molecule
    const int *bond_off;        // M + 1
    const int2 *bonds;
    const int *angle_off;
    const int4 *angles;         // (i, j, k, -)
    const int *tors_off;
    const int4 *tors;           // (i, j, k, l)
    const int *ind_off;
    const int2 *ind;            // indirect-neighbour pairs
    const int *local_off, *local;
    const int *crank_off;
    const int4 *crank;          // (residue before, residue, residue after, -) in the caller's order
    const int *flexr_off, *flexr;  // flex residues
    const int *nb_off;          // per flex residue entry: first neighbour entry
    const int *br_off, *br;     // per neighbour entry: residues the flex move rotates (caller order)
    const double *tor_v, *tor_d;   // 20 x 20 x 4 torsion table (kcal/mol, radians)
    const double *tor_cd, *tor_sd;  // cos and sin of tor_d
    double *out_int;            // R * 5 raw internal terms (full energy)
    // delayed acceptance with the surrogate (da = 0: plain Metropolis)
    int da;
    int *s1;                    // R: stage one accepted the pending proposal
    double *s_dS;               // R: surrogate change of the pending proposal
    float *smat;                // R * M * M surrogate pair energies of the current state
    float *srow;                // R * M: surrogate row of the pending proposal
    float *frames;              // R * M * 9 body frames of the current state
    float *prop_frame;          // R * 9 body frame of the moved molecule
    const int2 *anchor;         // M: anchor residues (device indices)
    const float *comp;          // M * 20 residue-type fractions
    const float *sp;            // surrogate FP32 masters
    const unsigned short *swb;  // surrogate BF16 weights
    long long *s1_rej;          // R: stage-one rejections
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// 64-bit DPP move: lane l receives the value of the lane the GFX9 control
// CTRL names (row_shr:n = 0x110 + n, row_bcast:15 = 0x142, row_bcast:31 =
// 0x143), or 0 where that lane is out of its row or ROW_MASK disables l's row.
// With every row enabled the bound control supplies the 0 (no zeroing move
// before each DPP move); a disabled row keeps the old value, 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ long long dpp_b64(long long v)
{
    constexpr bool all_rows = ROW_MASK == 0xf;
    const int lo = __builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, all_rows);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, ROW_MASK, 0xf, all_rows);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ long long readlane_b64(long long v, int l)
{
    const int lo = __builtin_amdgcn_readlane((int)v, l), hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Inclusive prefix sums over the 64 lanes (every lane active): shifts by
// 1, 2, 4, 8 within rows of 16, then row 0's total into row 1 and row 2's
// into row 3 (row_bcast:15), then rows 0-1's total into rows 2-3
// (row_bcast:31).  Six DPP steps instead of twelve LDS-crossbar shuffles.
__device__ __forceinline__ long long wave_scan_i64(long long v)
{
    v += dpp_b64<0x111, 0xf>(v);
    v += dpp_b64<0x112, 0xf>(v);
    v += dpp_b64<0x114, 0xf>(v);
    v += dpp_b64<0x118, 0xf>(v);
    v += dpp_b64<0x142, 0xa>(v);
    v += dpp_b64<0x143, 0xc>(v);
    return v;
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v)
{
    return __longlong_as_double(dpp_b64<CTRL, ROW_MASK>(__double_as_longlong(v)));
}

// Maximum over the 64 lanes of non-negative values (every lane active),
// returned in every lane: the same DPP steps with max.
__device__ __forceinline__ float wave_max_nonneg(float v)
{
#define CG_DPP_MAX(CTRL, RM) v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, RM, 0xf, false)))
    CG_DPP_MAX(0x111, 0xf);
    CG_DPP_MAX(0x112, 0xf);
    CG_DPP_MAX(0x114, 0xf);
    CG_DPP_MAX(0x118, 0xf);
    CG_DPP_MAX(0x142, 0xa);
    CG_DPP_MAX(0x143, 0xc);
#undef CG_DPP_MAX
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), kWave - 1));
}

// Sum over the 64 lanes (every lane active), returned in every lane: the
// DPP scan's last lane, read back.  A fixed summation order, so the result
// is reproducible.
__device__ __forceinline__ double wave_total(double v)
{
    v += dpp_f64<0x111, 0xf>(v);
    v += dpp_f64<0x112, 0xf>(v);
    v += dpp_f64<0x114, 0xf>(v);
    v += dpp_f64<0x118, 0xf>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    return __longlong_as_double(readlane_b64(__double_as_longlong(v), kWave - 1));
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float v)
{
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Loads the 20x20 table into LDS with the row/column layout used by the
// inner loop: byte offset (type_j << 9) | (type_i << 4).
__device__ __forceinline__ void load_table(const Coef *__restrict__ g, char *lds)
{
    for (int e = threadIdx.x; e < CGPPD_AA_COUNT * CGPPD_AA_COUNT; e += kWG) {
        const int tj = e / CGPPD_AA_COUNT, ti = e % CGPPD_AA_COUNT;
        *reinterpret_cast<Coef *>(lds + ((tj << kRowShift) | (ti << kColShift))) = g[e];
    }
}

// One pair term for this lane's residue i against the wave-uniform j record
// rj, with this lane's table entry cf for the (i type, j type) pair.
// MINIMG: apply the per-axis minimum image; otherwise the segment's single
// image shift is already folded into (xi, yi, zi).  CHARGED: the j residue
// carries charge, so the Debye-Hueckel term is evaluated.
template <bool MINIMG, bool CHARGED>
__device__ __forceinline__ void pair_term(const float4 rj, float xi, float yi, float zi, const Coef &cf, float box,
                                          float inv_box, float &lj, float &dh)
{
    float dx = xi - rj.x;
    float dy = yi - rj.y;
    float dz = zi - rj.z;
    if (MINIMG) {
        dx = fmaf(-box, rintf(dx * inv_box), dx);
        dy = fmaf(-box, rintf(dy * inv_box), dy);
        dz = fmaf(-box, rintf(dz * inv_box), dz);
    }
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float s2 = cf.sig2 * (rinv * rinv);
    const float s6 = s2 * s2 * s2;
    const float att = fmaf(cf.A, s6, -cf.A) * s6;
    const float rep = fmaf(cf.A, -0.5f, -att);
    lj = r2 < cf.rc2 ? rep : att;
    if (CHARGED) {
        const float r = r2 * rinv;
        dh = cf.qq * __builtin_amdgcn_exp2f(r * kDHExp2) * rinv;
    } else {
        dh = 0.0f;
    }
}

// A run [a, b) of j residues that share one type (wave-uniform), against
// this lane's residue i.  Terms are summed plainly in groups of four and the
// group sums enter a Kahan-compensated accumulator.  SEPARATE: LJ and DH
// accumulated apart (raw table); otherwise one total (scaled table).
template <bool SEPARATE, bool MINIMG, bool CHARGED>
__device__ __forceinline__ void pair_run(const float4 *__restrict__ jrec, int a, int b, float xi, float yi, float zi,
                                         const Coef &cf, float box, float inv_box, float &s0, float &c0, float &s1,
                                         float &c1)
{
    int j = a;
    for (; j + 4 <= b; j += 4) {
        float g0 = 0.0f, g1 = 0.0f;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float lj, dh;
            pair_term<MINIMG, CHARGED>(jrec[j + u], xi, yi, zi, cf, box, inv_box, lj, dh);
            if (SEPARATE) {
                g0 += lj;
                if (CHARGED) g1 += dh;
            } else {
                g0 += CHARGED ? lj + dh : lj;
            }
        }
        kahan_add(s0, c0, g0);
        if (SEPARATE && CHARGED) kahan_add(s1, c1, g1);
    }
    // the remaining 0-3 terms form one more group
    const int n = b - j;
    if (n > 0) {
        float g0 = 0.0f, g1 = 0.0f;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            if (u < n) {
                float lj, dh;
                pair_term<MINIMG, CHARGED>(jrec[j + u], xi, yi, zi, cf, box, inv_box, lj, dh);
                if (SEPARATE) {
                    g0 += lj;
                    if (CHARGED) g1 += dh;
                } else {
                    g0 += CHARGED ? lj + dh : lj;
                }
            }
        }
        kahan_add(s0, c0, g0);
        if (SEPARATE && CHARGED) kahan_add(s1, c1, g1);
    }
}

// Inner loop over the j residues [js, je) of one molecule (wave-uniform,
// indices relative to the staged chunk starting at jbase).  On the device
// each molecule's residues are ordered charged-first, then by type, so the
// range splits into constant-type runs: the lane's table entry is loaded
// once per run and uncharged runs skip the Debye-Hueckel term.
template <bool SEPARATE, bool MINIMG>
__device__ __forceinline__ void pair_segment(const float4 *__restrict__ jrec, int js, int je, int jbase,
                                             const PairParams &p, float xi, float yi, float zi, uint32_t lane_off,
                                             const char *__restrict__ tab, float &s0, float &c0, float &s1,
                                             float &c1)
{
    int run = p.run_of[jbase + js];
    int a = js;
    while (a < je) {
        const int b = min(je, p.run_start[run + 1] - jbase);
        const uint32_t code = p.run_code[run];
        const Coef cf = *reinterpret_cast<const Coef *>(tab + ((code & kRowMask) | lane_off));
        if (code & kChargedBit)
            pair_run<SEPARATE, MINIMG, true>(jrec, a, b, xi, yi, zi, cf, p.box, p.inv_box, s0, c0, s1, c1);
        else
            pair_run<SEPARATE, MINIMG, false>(jrec, a, b, xi, yi, zi, cf, p.box, p.inv_box, s0, c0, s1, c1);
        a = b;
        ++run;
    }
}

// Minimum-image shift shared by a whole segment.  With the two molecules'
// centres ci, cj and bounding radii (max |relative position|) summing to
// rsum, if every axis satisfies |d - L k| + rsum < L/2 (d = ci - cj,
// k = rint(d/L)), then every residue pair of the two molecules has the same
// minimum image k and the per-pair rint can be skipped.  Returns the shift
// L k to subtract from the i coordinates, or false when the segment needs the
// per-pair minimum image.
__device__ __forceinline__ bool segment_shift(float4 ci, float4 cj, float rsum, float box, float inv_box, float &sx,
                                              float &sy, float &sz)
{
    const float lim = 0.5f * box - (rsum * 1.0001f + 0.01f);
    const float kx = rintf((ci.x - cj.x) * inv_box), ky = rintf((ci.y - cj.y) * inv_box),
                kz = rintf((ci.z - cj.z) * inv_box);
    const float dx = fmaf(-box, kx, ci.x - cj.x), dy = fmaf(-box, ky, ci.y - cj.y), dz = fmaf(-box, kz, ci.z - cj.z);
    sx = box * kx;
    sy = box * ky;
    sz = box * kz;
    return fabsf(dx) < lim && fabsf(dy) < lim && fabsf(dz) < lim;
}

// Repulsive crowding (the reference's compile-time REPULSIVE_CROWDING,
// src/definitions.h:18; src/Replica.cpp:416-483, src/CudaFunctions.cu:335-338,
// 492-500): a residue pair with a crowder molecule contributes (6/r)^12 to LJ
// for r < 15 A (src/definitions.h:38) and no Debye-Hueckel term.  With the
// minimum-image differences (dx, dy, dz); raw LJ units.
constexpr float kCrowdCut2 = 225.0f;
constexpr float kSix12 = 2176782336.0f;  // 6^12, exact in float
__device__ __forceinline__ float crowd_term(float dx, float dy, float dz)
{
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float i2 = __builtin_amdgcn_rcpf(r2);
    const float u = i2 * i2 * i2;
    return r2 < kCrowdCut2 ? kSix12 * (u * u) : 0.0f;
}

__device__ __forceinline__ float min_image(float d, float box, float inv_box)
{
    return fmaf(-box, rintf(d * inv_box), d);
}

// A segment whose pairs all involve a crowder: this lane's residue pi
// against the j residues [js, je), LJ only (kcal/mol unless SEPARATE).
template <bool SEPARATE>
__device__ __forceinline__ void crowd_segment(const float4 *__restrict__ jrec, int js, int je, float4 pi,
                                              const PairParams &p, float &s0, float &c0)
{
    const float scale = SEPARATE ? 1.0f : p.rt_to_kcal;
    for (int j = js; j < je; ++j) {
        const float4 rj = jrec[j];
        const float e = crowd_term(min_image(pi.x - rj.x, p.box, p.inv_box), min_image(pi.y - rj.y, p.box, p.inv_box),
                                   min_image(pi.z - rj.z, p.box, p.inv_box));
        kahan_add(s0, c0, e * scale);
    }
}

// A segment (i-block of one molecule against the j residues [js, je) of
// molecule J): the shifted fast path when the bounding spheres allow it.
// crowd: either molecule is a crowder under repulsive crowding.
template <bool SEPARATE>
__device__ __forceinline__ void segment(const float4 *__restrict__ jrec, int js, int je, int jbase, float4 pi,
                                        uint32_t lane_off, const char *__restrict__ tab, const PairParams &p,
                                        float4 ci, float ri, float4 cj, float rj, bool crowd, float &s0, float &c0,
                                        float &s1, float &c1)
{
    if (crowd) {
        crowd_segment<SEPARATE>(jrec, js, je, pi, p, s0, c0);
        return;
    }
    float sx, sy, sz;
    if (p.fast_image && segment_shift(ci, cj, ri + rj, p.box, p.inv_box, sx, sy, sz))
        pair_segment<SEPARATE, false>(jrec, js, je, jbase, p, pi.x - sx, pi.y - sy, pi.z - sz, lane_off, tab, s0, c0,
                                      s1, c1);
    else
        pair_segment<SEPARATE, true>(jrec, js, je, jbase, p, pi.x, pi.y, pi.z, lane_off, tab, s0, c0, s1, c1);
}

// ---------------------------------------------------------------------------
// Flexible molecules: intra-molecular pairs and bonded terms
// ---------------------------------------------------------------------------
__device__ __forceinline__ int mulhi_n(uint32_t w, int n) { return (int)(((uint64_t)w * (uint64_t)n) >> 32); }

// The GPU kernel's exclusion of an intra-molecular pair
// (src/CudaFunctions.cu:447-476): same rigid domain, same segment bond, or
// the same chain with caller indices less than 4 apart.
__device__ __forceinline__ bool intra_excluded(int4 a, int4 b)
{
    return (a.y != 0 && a.y == b.y) || (a.z != 0 && a.z == b.z) || (a.x == b.x && abs(a.w - b.w) < 4);
}

__device__ __forceinline__ uint32_t lane_col(float4 pi) { return (__float_as_uint(pi.w) >> kRowShift & 0x1Fu) << kColShift; }

// One FP32 pair term between residues pi and pj (per-pair minimum image).
__device__ __forceinline__ void pair_any(float4 pi, float4 pj, const char *__restrict__ tab, const PairParams &p,
                                         float &lj, float &dh)
{
    const Coef cf = *reinterpret_cast<const Coef *>(tab + ((__float_as_uint(pj.w) & kRowMask) | lane_col(pi)));
    pair_term<true, true>(pj, pi.x, pi.y, pi.z, cf, p.box, p.inv_box, lj, dh);
}

// pair_any with the coefficients read from a global table (entry tj * 20 + ti).
__device__ __forceinline__ void pair_global(float4 pi, float4 pj, const Coef *__restrict__ g, const PairParams &p,
                                            float &lj, float &dh)
{
    const uint32_t ti = __float_as_uint(pi.w) >> kRowShift & 0x1Fu, tj = __float_as_uint(pj.w) >> kRowShift & 0x1Fu;
    const Coef cf = g[tj * CGPPD_AA_COUNT + ti];
    pair_term<true, true>(pj, pi.x, pi.y, pi.z, cf, p.box, p.inv_box, lj, dh);
}

// Intra-molecular item: this lane's residue a = ib*64 + lane of a molecule
// of n residues (molecule-local indices, coordinates src[], tags tg[])
// against the residues b > a of chunk jc (256), wave w taking 64 of them.
// Kahan-compensated FP32 sums of LJ and DH.
__device__ __forceinline__ void intra_item(const float4 *__restrict__ src, const int4 *__restrict__ tg, int n, int ib,
                                           int jc, const PairParams &p, const char *__restrict__ tab, float4 *jrec,
                                           int4 *jtag, float &s0, float &c0, float &s1, float &c1, bool &valid)
{
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    __syncthreads();  // previous item finished reading the staging arrays
    {
        const int j = jc * kIntraChunk + threadIdx.x;
        if (threadIdx.x < kIntraChunk && j < n) {
            jrec[threadIdx.x] = src[j];
            jtag[threadIdx.x] = tg[j];
        }
    }
    const int a = ib * kIBlock + lane;
    valid = a < n;
    const float4 pa = src[valid ? a : 0];
    const int4 ta = tg[valid ? a : 0];
    __syncthreads();
    constexpr int kPerWave = kIntraChunk / kWavesPerWG;
    const int j0 = jc * kIntraChunk + wave * kPerWave;
    // a wave whose j residues all precede its i block has no pair b > a
    const int j1 = j0 + kPerWave - 1 <= ib * kIBlock ? j0 : min(n, j0 + kPerWave);
    // four j residues at a time: their records, tags and table entries are
    // read before any term is formed (the terms enter the sums in j order)
    const int jb = jc * kIntraChunk;
    const uint32_t col = lane_col(pa);
    int b = j0;
    for (; b + 4 <= j1; b += 4) {
        float4 pb[4];
        int4 tb[4];
        Coef cf[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            pb[u] = jrec[b + u - jb];
            tb[u] = jtag[b + u - jb];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) cf[u] = *reinterpret_cast<const Coef *>(tab + ((__float_as_uint(pb[u].w) & kRowMask) | col));
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float lj, dh;
            pair_term<true, true>(pb[u], pa.x, pa.y, pa.z, cf[u], p.box, p.inv_box, lj, dh);
            if (valid && b + u > a && !intra_excluded(ta, tb[u])) {
                kahan_add(s0, c0, lj);
                kahan_add(s1, c1, dh);
            }
        }
    }
    for (; b < j1; ++b) {
        const float4 pb = jrec[b - jb];
        const int4 tb = jtag[b - jb];
        float lj, dh;
        pair_any(pa, pb, tab, p, lj, dh);
        if (valid && b > a && !intra_excluded(ta, tb)) {
            kahan_add(s0, c0, lj);
            kahan_add(s1, c1, dh);
        }
    }
}

// Internal-energy change item of a flexible molecule whose proposal moves
// one residue a (molecule-local; proposed position src[a], current cur[a]):
// thread t takes residue b = jc * 256 + t of the molecule, b != a, and
// sums e(new_a, b) - e(cur_a, b) over the pairs the GPU kernel's exclusions
// keep (Kahan-compensated FP32 per thread; the caller reduces the waves).
__device__ __forceinline__ void moved_item(const float4 *__restrict__ src, const float4 *__restrict__ cur,
                                           const int4 *__restrict__ tg, int n, int a, int jc, const PairParams &p,
                                           const char *__restrict__ tab, float &s0, float &c0)
{
    const int b = jc * kJChunk + threadIdx.x;
    if (b >= n || b == a) return;
    const int4 ta = tg[a], tb = tg[b];
    if (intra_excluded(ta, tb)) return;
    const float4 pb = src[b];
    float ljn, dhn, ljo, dho;
    pair_any(src[a], pb, tab, p, ljn, dhn);
    pair_any(cur[a], pb, tab, p, ljo, dho);
    kahan_add(s0, c0, (ljn + dhn) - (ljo + dho));
}

// src/Quaternion.cpp:95-112, double arithmetic on a float vector.
__device__ __forceinline__ float3 qrot(const double q[4], float vx_, float vy_, float vz_)
{
#pragma clang fp contract(off)
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double w2 = w * w, x2 = x * x, y2 = y * y, z2 = z * z;
    const double wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
    const double vx = vx_, vy = vy_, vz = vz_;
    return make_float3((float)(vx * (w2 + x2 - y2 - z2) + (vy * (xy - wz) + vz * (wy + xz)) * 2.0),
                       (float)((vx * (wz + xy) + vz * (yz - wx)) * 2.0 + vy * (w2 - x2 + y2 - z2)),
                       (float)((vx * (xz - wy) + vy * (yz + wx)) * 2.0 + vz * (w2 - x2 - y2 + z2)));
}

__device__ __forceinline__ float fmag(float x, float y, float z)
{
#pragma clang fp contract(off)
    return sqrtf(x * x + y * y + z * z);
}

// Bonded terms of molecule m (whose residues start at device index i0;
// coordinates src[] molecule-local) with the reference's float vectors:
// bond src/Potential.cpp:45-55, angle :57-71 (src/vector3f.cpp:137-143),
// torsion :73-98, and the indirect-neighbour pair terms that
// Molecule::E(false) subtracts (src/Molecule.cpp:669-683).  Every thread
// takes a strided share; t[5] = indirect LJ and DH (negated), bond sum,
// sum of log(angle term), torsion sum.
// One bonded or indirect-neighbour term of each kind, from its residues'
// coordinates.
__device__ __forceinline__ double bond_term(float4 a, float4 b, float box)
{
#pragma clang fp contract(off)
    float X = b.x - a.x, Y = b.y - a.y, Z = b.z - a.z;
    X = X - box * roundf(X / box);
    Y = Y - box * roundf(Y / box);
    Z = Z - box * roundf(Z / box);
    const double r = (double)sqrtf(X * X + Y * Y + Z * Z);
    return (r - (double)3.81f) * (r - (double)3.81f);
}

__device__ __forceinline__ double angle_term(float4 a, float4 b, float4 c)
{
#pragma clang fp contract(off)
    const float abx = a.x - b.x, aby = a.y - b.y, abz = a.z - b.z;
    const float cbx = c.x - b.x, cby = c.y - b.y, cbz = c.z - b.z;
    const double cmag = (double)(fmag(cbx, cby, cbz) * fmag(abx, aby, abz));
    float th = 0.0f;
    if (!(cmag < 1e-26)) th = (float)acos((double)(abx * cbx + aby * cby + abz * cbz) / cmag);
    const double th_d = (double)th;
    const double da = th_d - (double)1.6f, db = th_d - (double)2.27f;
    // log(e^u + e^v) = max(u, v) + log1p(e^-|u - v|): one exp fewer, the
    // same FP64 value to rounding
    const double u = -(double)0.1f * ((double)106.4f * da * da + (double)4.3f);
    const double v = -(double)0.1f * ((double)26.3f * db * db);
    return fmax(u, v) + log1p(exp(-fabs(u - v)));
}

__device__ __forceinline__ double torsion_term(float4 a, float4 b, float4 c, float4 d, const PairParams &p)
{
#pragma clang fp contract(off)
    const float b1x = b.x - a.x, b1y = b.y - a.y, b1z = b.z - a.z;
    const float b2x = c.x - b.x, b2y = c.y - b.y, b2z = c.z - b.z;
    const float b3x = d.x - c.x, b3y = d.y - c.y, b3z = d.z - c.z;
    const float x23 = b2y * b3z - b2z * b3y, y23 = b2z * b3x - b2x * b3z, z23 = b2x * b3y - b2y * b3x;
    const float x12 = b1y * b2z - b1z * b2y, y12 = b1z * b2x - b1x * b2z, z12 = b1x * b2y - b1y * b2x;
    const float yy = fmag(b2x, b2y, b2z) * (b1x * x23 + b1y * y23 + b1z * z23);
    const float xx = x12 * x23 + y12 * y23 + z12 * z23;
    const double phi = (double)atan2f(yy, xx);
    const int tj = (int)(__float_as_uint(b.w) >> kRowShift & 0x1Fu);
    const int tk = (int)(__float_as_uint(c.w) >> kRowShift & 0x1Fu);
    const int e = (tj * 20 + tk) * 4;
    const double *V = p.tor_v + e, *CD = p.tor_cd + e, *SD = p.tor_sd + e;
    // cos(n phi - d_n) = cos(n phi) cos d_n + sin(n phi) sin d_n, with
    // cos(n phi) and sin(n phi) from one sincos by the angle-addition
    // formulas (the same FP64 value as four cos calls, to rounding)
    double s1, c1;
    sincos(phi, &s1, &c1);
    const double c2 = c1 * c1 - s1 * s1, s2 = 2.0 * (s1 * c1);
    const double c3 = c2 * c1 - s2 * s1, s3 = s2 * c1 + c2 * s1;
    const double c4 = c2 * c2 - s2 * s2, s4 = 2.0 * (s2 * c2);
    return (1 + (c1 * CD[0] + s1 * SD[0])) * V[0] + (1 + (c2 * CD[1] + s2 * SD[1])) * V[1] +
           (1 + (c3 * CD[2] + s3 * SD[2])) * V[2] + (1 + (c4 * CD[3] + s4 * SD[3])) * V[3];
}

// Bonded terms of flexible molecule m and the indirect-neighbour pairs that
// Molecule::E(false) subtracts (src/Molecule.cpp:669-683), from the
// molecule-local coordinates src (whole workgroup, the terms spread over
// its threads).
// t[5] = indirect LJ and DH (negated), bond sum, sum of log(angle term),
// torsion sum, this thread's share.  With old non-null, every term is
// evaluated on both coordinate sets and t holds new minus old; with only
// >= 0 as well (the molecule-local index of the one residue whose
// coordinates differ between src and old), the terms without it are skipped.
// GLOBAL_TABLE: the indirect-neighbour pairs read the scaled table from
// global memory (kernels without the LDS table), else the LDS table tab.
constexpr int kBondedBatch = 4;

// The bonded terms of flexible molecule m in one index space, the costliest
// kinds first (torsions, angles, bonds, indirect pairs), each kind's range
// padded to whole waves so that a wave evaluates one kind at a time.
struct BondedSpace {
    int i0, to, ao, bo, io, nt, na, nb, qa, qb, qi, total;
    __device__ BondedSpace(const PairParams &p, int m)
    {
        i0 = p.mol_off[m];
        to = p.tors_off[m], ao = p.angle_off[m], bo = p.bond_off[m], io = p.ind_off[m];
        nt = p.tors_off[m + 1] - to, na = p.angle_off[m + 1] - ao;
        nb = p.bond_off[m + 1] - bo;
        const int ni = p.ind_off[m + 1] - io;
        qa = (nt + kWave - 1) & ~(kWave - 1), qb = qa + ((na + kWave - 1) & ~(kWave - 1));
        qi = qb + ((nb + kWave - 1) & ~(kWave - 1)), total = qi + ni;
    }
};

// The kind (0 torsion, 1 angle, 2 bond, 3 indirect pair, -1 none) and
// residues of the terms q0 + u * blockDim.x, u < kBondedBatch (unused
// residue slots name the molecule's first residue).
struct BondedBatch {
    int kind[kBondedBatch];
    int4 e[kBondedBatch];
};

__device__ __forceinline__ void bonded_batch(const PairParams &p, const BondedSpace &sp, int q0, BondedBatch &bt)
{
    const int i0 = sp.i0;
#pragma unroll
    for (int u = 0; u < kBondedBatch; ++u) {
        const int q = q0 + u * blockDim.x;
        bt.kind[u] = q >= sp.total ? -1
                     : q < sp.qa   ? (q < sp.nt ? 0 : -1)
                     : q < sp.qb   ? (q - sp.qa < sp.na ? 1 : -1)
                     : q < sp.qi   ? (q - sp.qb < sp.nb ? 2 : -1)
                                   : 3;
        bt.e[u] = make_int4(i0, i0, i0, i0);
        if (bt.kind[u] == 0) {
            bt.e[u] = p.tors[sp.to + q];
        } else if (bt.kind[u] == 1) {
            const int4 a = p.angles[sp.ao + q - sp.qa];
            bt.e[u] = make_int4(a.x, a.y, a.z, i0);
        } else if (bt.kind[u] >= 2) {
            const int2 b = bt.kind[u] == 2 ? p.bonds[sp.bo + q - sp.qb] : p.ind[sp.io + q - sp.qi];
            bt.e[u] = make_int4(b.x, b.y, i0, i0);
        }
    }
}

template <bool GLOBAL_TABLE>
__device__ void bonded_terms(const float4 *__restrict__ src, int i0, int m, const PairParams &p,
                             const char *__restrict__ tab, double t[5], const float4 *__restrict__ old = nullptr,
                             int only = -1)
{
#pragma clang fp contract(off)
    for (int k = 0; k < 5; ++k) t[k] = 0.0;
    // thread t takes entries t, t + block, ..., kBondedBatch at a time
    // (their indices, then their coordinates, then the arithmetic)
    const BondedSpace sp(p, m);
    const int g = i0 + only;  // device index of the only moved residue
    constexpr int kBatch = kBondedBatch;
    for (int q0 = threadIdx.x; q0 < sp.total; q0 += kBatch * blockDim.x) {
        BondedBatch bt;
        bonded_batch(p, sp, q0, bt);
        int kind[kBatch];
        int4 e[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            kind[u] = bt.kind[u];
            e[u] = bt.e[u];
        }
        if (only >= 0) {
            // one residue moved and nothing else did: the other terms'
            // differences are exactly zero
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const bool has = e[u].x == g || e[u].y == g || (kind[u] <= 1 && e[u].z == g) || (kind[u] == 0 && e[u].w == g);
                if (!has) kind[u] = -1;
            }
        }
        float4 A[kBatch][4], O[kBatch][4];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (kind[u] < 0) continue;
            A[u][0] = src[e[u].x - i0];
            A[u][1] = src[e[u].y - i0];
            A[u][2] = src[e[u].z - i0];
            A[u][3] = src[e[u].w - i0];
            if (old) {
                O[u][0] = old[e[u].x - i0];
                O[u][1] = old[e[u].y - i0];
                O[u][2] = old[e[u].z - i0];
                O[u][3] = old[e[u].w - i0];
            }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (kind[u] == 0) {
                t[4] += torsion_term(A[u][0], A[u][1], A[u][2], A[u][3], p);
                if (old) t[4] -= torsion_term(O[u][0], O[u][1], O[u][2], O[u][3], p);
            } else if (kind[u] == 1) {
                t[3] += angle_term(A[u][0], A[u][1], A[u][2]);
                if (old) t[3] -= angle_term(O[u][0], O[u][1], O[u][2]);
            } else if (kind[u] == 2) {
                t[2] += bond_term(A[u][0], A[u][1], p.box);
                if (old) t[2] -= bond_term(O[u][0], O[u][1], p.box);
            } else if (kind[u] == 3) {
                float lj, dh;
                if (GLOBAL_TABLE)
                    pair_global(A[u][0], A[u][1], p.table_scaled, p, lj, dh);
                else
                    pair_any(A[u][0], A[u][1], tab, p, lj, dh);
                t[0] -= (double)lj;
                t[1] -= (double)dh;
                if (old) {
                    if (GLOBAL_TABLE)
                        pair_global(O[u][0], O[u][1], p.table_scaled, p, lj, dh);
                    else
                        pair_any(O[u][0], O[u][1], tab, p, lj, dh);
                    t[0] += (double)lj;
                    t[1] += (double)dh;
                }
            }
        }
    }
}

// kcal/mol of bonded_terms' vector when the pair table is scaled (the
// indirect-neighbour terms are then already kcal/mol): src/Potential.cpp:203-210.
__device__ __forceinline__ double bonded_kcal(const double t[5])
{
    return t[0] + t[1] + t[2] * (0.5 * 378.0) + t[3] * -(double)10.0f + t[4];
}

// Maps a persistent workgroup's k-th item to (replica, item-in-replica).
// With xcd_map, workgroup b runs on XCD b % 8 and only takes items of
// replicas r with r % 8 == b % 8, so a replica's coordinates live in one L2.
struct ItemIter {
    int k, step, count, per_rep, xcd, base;
    __device__ ItemIter(const PairParams &p, int per_rep_)
    {
        per_rep = per_rep_;
        const int G = gridDim.x, b = blockIdx.x;
        if (p.xcd_map) {
            xcd = b & 7;
            k = b >> 3;
            step = G >> 3;
            count = (p.rep_count >> 3) * per_rep;
        } else {
            xcd = -1;
            k = b;
            step = G;
            count = p.rep_count * per_rep;
        }
        base = p.rep_begin;
    }
    __device__ bool valid() const { return k < count; }
    __device__ void next() { k += step; }
    __device__ int replica() const { return base + (xcd < 0 ? k / per_rep : xcd + 8 * (k / per_rep)); }
    __device__ int item() const { return k % per_rep; }
};

// ---------------------------------------------------------------------------
// Full energy: rows I of every replica in [rep_begin, rep_begin+rep_count)
// against all J > I (J < n_active).  SCALED: single total in kcal/mol into
// the MC matrix; otherwise raw LJ/DH totals (+ optional matrix).
// ---------------------------------------------------------------------------
template <bool SCALED>
__global__ __launch_bounds__(kWG) void full_energy_kernel(PairParams p)
{
    __shared__ __attribute__((aligned(16))) char tab[kTableBytes];
    __shared__ float4 jrec[kJChunk];
    load_table(p.table, tab);

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int per_rep = p.n_active * p.ib_max * p.jc_count;
    for (ItemIter it(p, per_rep); it.valid(); it.next()) {
        const int r = it.replica();
        int q = it.item();
        const int jc = q % p.jc_count;
        q /= p.jc_count;
        const int ib = q % p.ib_max;
        const int I = q / p.ib_max;
        const int i_beg = p.mol_off[I], i_end = p.mol_off[I + 1];
        const int j_lo = i_end, j_hi = p.mol_off[p.n_active];
        const int j0 = max(jc * kJChunk, j_lo), j1 = min((jc + 1) * kJChunk, j_hi);
        if (i_beg + ib * kIBlock >= i_end || j0 >= j1) continue;  // uniform

        const float4 *pos = p.pos + (size_t)r * p.stride;
        __syncthreads();  // previous item finished reading jrec
        {
            const int j = jc * kJChunk + threadIdx.x;
            if (j < p.n_res) jrec[threadIdx.x] = pos[j];
        }
        const int i = i_beg + ib * kIBlock + lane;
        const bool valid = i < i_end;
        const float4 pi = pos[valid ? i : i_beg];
        const uint32_t lane_off = (__float_as_uint(pi.w) >> kRowShift & 0x1Fu) << kColShift;
        const float4 ci = p.center[(size_t)r * p.n_mol + I];
        const float ri = p.radius[(size_t)r * p.n_mol + I];
        __syncthreads();

        int js = max(j0, jc * kJChunk + wave * kWave);
        const int je = min(j1, jc * kJChunk + (wave + 1) * kWave);
        while (js < je) {
            const int J = p.mol_of[js];
            const int se = min(je, p.mol_off[J + 1]);
            float s0 = 0.f, c0 = 0.f, s1 = 0.f, c1 = 0.f;
            segment<!SCALED>(jrec, js - jc * kJChunk, se - jc * kJChunk, jc * kJChunk, pi, lane_off, tab, p, ci, ri,
                             p.center[(size_t)r * p.n_mol + J], p.radius[(size_t)r * p.n_mol + J],
                             J >= p.crowd_mol, s0, c0, s1, c1);
            double v0 = valid ? (double)s0 - (double)c0 : 0.0;
            double v1 = valid ? (double)s1 - (double)c1 : 0.0;
            v0 = wave_sum(v0);
            if (!SCALED) v1 = wave_sum(v1);
            if (lane == 0) {
                const size_t mm = (size_t)p.n_mol * p.n_mol;
                if (SCALED) {
                    atomicAdd(p.emat + r * mm + (size_t)I * p.n_mol + J, v0);
                } else {
                    atomicAdd(p.out_lj + r, v0);
                    atomicAdd(p.out_dh + r, v1);
                    if (p.out_mat)
                        atomicAdd(p.out_mat + r * mm + (size_t)I * p.n_mol + J,
                                  v0 * (double)p.rt_to_kcal + v1 * (double)p.dh_factor);
                }
            }
            js = se;
        }
    }
}


// Full energy of flexible molecules: the intra-molecular pairs the GPU
// kernel of the reference counts (src/CudaFunctions.cu:447-476) and the
// bonded terms (src/Molecule.cpp:636-699) of every flexible molecule of the
// replicas [rep_begin, rep_begin + rep_count).  Items per replica: for each
// flexible molecule, its (i-block, j-chunk) pairs then one bonded item
// (flex_item_off).  SCALED: kcal/mol into the MC matrix diagonal, with the
// indirect neighbours subtracted and the bonded terms added; otherwise raw
// LJ/DH into out_lj/out_dh (EonDevice), the raw internal terms into out_int
// and the intra-molecular total into out_mat's diagonal.
template <bool SCALED>
__global__ __launch_bounds__(kWG) void intra_energy_kernel(PairParams p)
{
    __shared__ __attribute__((aligned(16))) char tab[kTableBytes];
    __shared__ float4 jrec[kJChunk];
    __shared__ int4 jtag[kJChunk];
    load_table(p.table, tab);
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int per_rep = p.flex_item_off[p.n_flex];
    const size_t mm = (size_t)p.n_mol * p.n_mol;
    for (ItemIter it(p, per_rep); it.valid(); it.next()) {
        const int r = it.replica();
        const int q = it.item();
        int f = 0;
        while (q >= p.flex_item_off[f + 1]) ++f;
        const int m = p.flex_mols[f];
        const int i0 = p.mol_off[m], n = p.mol_off[m + 1] - i0;
        const int nib = (n + kIBlock - 1) / kIBlock, njc = intra_chunks(n);
        const float4 *src = p.pos + (size_t)r * p.stride + i0;
        const int local = q - p.flex_item_off[f];
        double *diag = p.emat + r * mm + (size_t)m * p.n_mol + m;
        if (local < nib * njc) {
            float s0 = 0.f, c0 = 0.f, s1 = 0.f, c1 = 0.f;
            bool valid;
            intra_item(src, p.rtag + i0, n, local / njc, local % njc, p, tab, jrec, jtag, s0, c0, s1, c1, valid);
            double v0 = valid ? (double)s0 - (double)c0 : 0.0, v1 = valid ? (double)s1 - (double)c1 : 0.0;
            v0 = wave_sum(v0);
            v1 = wave_sum(v1);
            if (lane == 0) {
                if (SCALED) {
                    atomicAdd(diag, v0 + v1);
                } else {
                    atomicAdd(p.out_lj + r, v0);
                    atomicAdd(p.out_dh + r, v1);
                    if (p.out_mat)
                        atomicAdd(p.out_mat + r * mm + (size_t)m * p.n_mol + m,
                                  v0 * (double)p.rt_to_kcal + v1 * (double)p.dh_factor);
                }
            }
        } else {
            double t[5];
            bonded_terms<false>(src, i0, m, p, tab, t);
            if (SCALED) {
                const double v = wave_sum(bonded_kcal(t));
                if (lane == 0) atomicAdd(diag, v);
            } else {
                for (int k = 0; k < 5; ++k) t[k] = wave_sum(t[k]);
                if (lane == 0)
                    for (int k = 0; k < 5; ++k) atomicAdd(p.out_int + (size_t)r * 5 + k, t[k]);
            }
        }
    }
}

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) f4_t const_f4;

// A wave-uniform record of the proposal through the scalar cache (the
// proposal is not written while the pair kernels run)
__device__ __forceinline__ f4_t scalar_rec(const float4 *base, int i)
{
    return ((const_f4 *)base)[i];
}

// Wave units of a flexible molecule's intra-molecular pairs in the MC pair
// kernel: unit (ib, jr) is the i block ib (64 residues, one per lane)
// against the kUnitJ residues of j range jr.  A range wholly before the i
// block (jr < kUnitRatio ib) holds no pair b > a and is not listed; the
// units run i block by i block, j ranges ascending.
constexpr int kUnitJ = 32;
constexpr int kUnitRatio = kIBlock / kUnitJ;  // j ranges per i block
static_assert(kIBlock == kUnitRatio * kUnitJ, "an i block spans whole j ranges");

__host__ __device__ inline int intra_units(int n)
{
    const int nib = (n + kIBlock - 1) / kIBlock, nr = (n + kUnitJ - 1) / kUnitJ;
    int u = 0;
    for (int ib = 0; ib < nib; ++ib) u += nr - kUnitRatio * ib > 0 ? nr - kUnitRatio * ib : 0;
    return u;
}

// (i block, j range) of unit u < intra_units(n)
__device__ __forceinline__ int2 intra_unit_of(int u, int n)
{
    const int nr = (n + kUnitJ - 1) / kUnitJ;
    int ib = 0;
    while (u >= nr - kUnitRatio * ib) {
        u -= nr - kUnitRatio * ib;
        ++ib;
    }
    return make_int2(ib, kUnitRatio * ib + u);
}

// Part [b0, b1) of a wave unit's j range (records and tags staged at
// jrec[b - j0], jtag[b - j0]) against this lane's residue a (record pa, tag
// ta, table column col), pairs b > a outside the exclusions; four j
// residues are read at a time and the terms enter the sums in b order.
// MINIMG: the per-pair minimum image.  CHARGED: the j residues carry
// charge (else the Debye-Hueckel term is zero and skipped).
template <bool MINIMG, bool CHARGED>
__device__ __forceinline__ void intra_run(int b0, int b1, int j0, int a, float4 pa, int4 ta, uint32_t col,
                                          const float4 *jrec, const int4 *jtag, const PairParams &p,
                                          const char *__restrict__ tab, float &s0, float &c0, float &s1, float &c1)
{
    int b = b0;
    for (; b + 4 <= b1; b += 4) {
        float4 rb[4];
        int4 gb[4];
        Coef cf[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            rb[u] = jrec[b + u - j0];
            gb[u] = jtag[b + u - j0];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            cf[u] = *reinterpret_cast<const Coef *>(tab + ((__float_as_uint(rb[u].w) & kRowMask) | col));
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float lj, dh;
            pair_term<MINIMG, CHARGED>(rb[u], pa.x, pa.y, pa.z, cf[u], p.box, p.inv_box, lj, dh);
            if (b + u > a && !intra_excluded(ta, gb[u])) {
                kahan_add(s0, c0, lj);
                if (CHARGED) kahan_add(s1, c1, dh);
            }
        }
    }
    for (; b < b1; ++b) {
        const float4 rb = jrec[b - j0];
        const int4 gb = jtag[b - j0];
        const Coef cf = *reinterpret_cast<const Coef *>(tab + ((__float_as_uint(rb.w) & kRowMask) | col));
        float lj, dh;
        pair_term<MINIMG, CHARGED>(rb, pa.x, pa.y, pa.z, cf, p.box, p.inv_box, lj, dh);
        if (b > a && !intra_excluded(ta, gb)) {
            kahan_add(s0, c0, lj);
            if (CHARGED) kahan_add(s1, c1, dh);
        }
    }
}

// One wave unit: this lane's residue a = ib * 64 + lane of a molecule of n
// residues (molecule-local coordinates src[], tags tg[]) against the
// residues b > a of j range jr, whose records and tags the caller staged in
// LDS, the first nq of them charged (the device order puts a molecule's
// charged residues first).  Kahan-compensated FP32 sums of LJ and DH.
// MINIMG: the per-pair minimum image; without it the plain differences
// (the caller's molecule spans less than half the box, so the minimum
// image is the plain difference).
template <bool MINIMG>
__device__ __forceinline__ void intra_unit(const float4 *__restrict__ src, const int4 *__restrict__ tg, int n, int ib,
                                           int jr, int nq, const float4 *jrec, const int4 *jtag, const PairParams &p,
                                           const char *__restrict__ tab, float &s0, float &c0, float &s1, float &c1,
                                           bool &valid)
{
    const int a0 = ib * kIBlock + (threadIdx.x & (kWave - 1));
    valid = a0 < n;
    const float4 pa = src[valid ? a0 : 0];
    const int4 ta = tg[valid ? a0 : 0];
    const int a = valid ? a0 : 0x7fffffff;  // (no b > a for lanes past the molecule)
    const uint32_t col = lane_col(pa);
    const int j0 = jr * kUnitJ, j1 = min(n, j0 + kUnitJ), jq = j0 + nq;
    intra_run<MINIMG, true>(j0, jq, j0, a, pa, ta, col, jrec, jtag, p, tab, s0, c0, s1, c1);
    intra_run<MINIMG, false>(jq, j1, j0, a, pa, ta, col, jrec, jtag, p, tab, s0, c0, s1, c1);
}

// Inter-molecular items of the i-lane kernel for a moved molecule of n
// residues in nib i blocks: none when the molecule is alone in the system.
__device__ __forceinline__ int ilane_inter_items(const PairParams &p, int nib, int n)
{
    return p.n_res > n ? nib * p.jc_count : 0;
}

// MC step, phase 1 (pair kernel): the delta-E row of every replica's moved
// molecule against all other molecules, FP32 pair terms only.  The proposed
// residue positions of the moved molecule were laid out by the previous
// phase 2 (prop_pos, device order, w = residue code), so this kernel carries
// no double-precision move code.  Work items are the active (replica,
// i-block, 256-residue j-chunk) triples, listed by prefix sum over the
// replicas' i-block counts and split into equal contiguous ranges -- one per
// workgroup, workgroups ordered by XCD so a replica's items share an L2.
// FLEX: a flexible moved molecule adds its intra-molecular items, summed
// into the row's diagonal entry: after a move of all its residues, four
// wave units (intra_unit) per item; after a one-residue move, the moved
// residue's pairs (moved_item).
template <bool FLEX>
__global__ __launch_bounds__(kWG) void mc_pair_kernel(PairParams p)
{
    __shared__ __attribute__((aligned(16))) char tab[kTableBytes];
    __shared__ float4 jrec[kJChunk];
    __shared__ int4 jtag[FLEX ? kWavesPerWG * kUnitJ : 1];
    __shared__ int item_off[CGPPD_MAX_REPLICAS + 1];
    // set-up and item staging at high wave priority, the pair loops at
    // normal priority (see mc_row_kernel)
    __builtin_amdgcn_s_setprio(2);
    const int R = p.rep_count, JC = p.jc_count;
    // the set-up's loads together: the table (two entries per thread) and
    // the first 64 replicas' item counts (recorded with each proposal; zero
    // for a proposal rejected by the surrogate's stage one)
    const int lane0 = threadIdx.x & (kWave - 1);
    const int e0 = threadIdx.x, e1 = min((int)threadIdx.x + kWG, CGPPD_AA_COUNT * CGPPD_AA_COUNT - 1);
    const Coef g0 = p.table[e0], g1 = p.table[e1];
    const int it0 = p.prop_items[min(lane0, R - 1)];
    *reinterpret_cast<Coef *>(tab + (((e0 / CGPPD_AA_COUNT) << kRowShift) | ((e0 % CGPPD_AA_COUNT) << kColShift))) = g0;
    *reinterpret_cast<Coef *>(tab + (((e1 / CGPPD_AA_COUNT) << kRowShift) | ((e1 % CGPPD_AA_COUNT) << kColShift))) = g1;
    // item offsets of the replicas: an inclusive DPP scan, 64 replicas at a
    // time; every wave scans (an unconditional use of the loads), wave 0
    // stores
    {
        auto items = [&](int r, int c) -> long long { return (r < R && !(p.da && !p.s1[r])) ? c : 0; };
        long long c = wave_scan_i64(items(lane0, it0));
        if (threadIdx.x < kWave && lane0 < R) item_off[lane0 + 1] = (int)c;
        long long carry = readlane_b64(c, kWave - 1);
        for (int base = kWave; base < R; base += kWave) {
            const int r = base + lane0;
            c = wave_scan_i64(items(r, p.prop_items[min(r, R - 1)]));
            if (threadIdx.x < kWave && r < R) item_off[r + 1] = (int)(carry + c);
            carry += readlane_b64(c, kWave - 1);
        }
        if (threadIdx.x == 0) item_off[0] = 0;
    }
    __syncthreads();
    const int total = item_off[R];
    const int G = gridDim.x, b = blockIdx.x;
    const int lb = (p.xcd_map && (G & 7) == 0) ? (b & 7) * (G >> 3) + (b >> 3) : b;
    const int k0 = (int)((long long)lb * total / G), k1 = (int)((long long)(lb + 1) * total / G);

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    // replica of the first item: the number of replicas that end at or
    // before it (item_off is non-decreasing), 64 at a time; then advance
    int r = 0;
    if (k0 < k1) {
        for (int base = 0; base < R; base += kWave) {
            const int k = base + lane;
            const unsigned long long before = __ballot(k < R && item_off[k + 1] <= k0);
            r = base + __popcll(before);
            if (before != ~0ull) break;
        }
    }
    for (int k = k0; k < k1; ++k) {
        while (k >= item_off[r + 1]) ++r;
        const int local = k - item_off[r];
        // (the proposal header holds the moved molecule, its first residue
        // and its size: one load level)
        const int4 hd = FLEX ? p.prop_hdr[r] : make_int4(0, 0, 0, 0);
        const int4 mv = FLEX ? p.prop_moved[r] : make_int4(0, 0, 0, 0);
        const int n_inter = FLEX ? ilane_inter_items(p, (hd.z + kIBlock - 1) / kIBlock, hd.z) : 0;
        if (FLEX && local >= n_inter) {
            const int m = hd.x, i0 = hd.y, n = hd.z;
            const int l2 = local - n_inter;
            const float4 *src = p.prop_pos + (size_t)r * p.max_n;
            double v = 0.0;
            // (the bonded terms were added to the diagonal when the proposal
            // was laid out)
            __builtin_amdgcn_s_setprio(0);
            if (mv.x == kMovedAll) {
                // the intra-molecular pairs of the proposal: this wave's
                // unit, its j range staged in the wave's part of jrec / jtag
                const int u = l2 * kWavesPerWG + wave;
                const bool has = u < intra_units(n);
                const int2 q = has ? intra_unit_of(u, n) : make_int2(0, 0);
                float4 *wrec = jrec + wave * kUnitJ;
                int4 *wtag = jtag + wave * kUnitJ;
                __builtin_amdgcn_s_setprio(2);
                __syncthreads();  // the previous item finished reading jrec
                int nq;
                {
                    const int j = q.y * kUnitJ + lane;
                    const bool mine = has && lane < kUnitJ && j < n;
                    float4 rj = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (mine) {
                        rj = src[j];
                        wrec[lane] = rj;
                        wtag[lane] = p.rtag[i0 + j];
                    }
                    // the range's charged residues (a prefix of it)
                    nq = __popcll(__ballot(mine && (__float_as_uint(rj.w) & kChargedBit)));
                }
                __syncthreads();
                if (!has) continue;
                __builtin_amdgcn_s_setprio(0);
                float s0 = 0.f, c0 = 0.f, s1 = 0.f, c1 = 0.f;
                bool valid;
                // every residue lies within the proposal's bounding radius of
                // its centre: below a quarter box (less a margin for the
                // rounding of centre plus relative position) no pair needs
                // an image
                if (4.0f * p.prop_radius[r] < 0.98f * p.box)
                    intra_unit<false>(src, p.rtag + i0, n, q.x, q.y, nq, wrec, wtag, p, tab, s0, c0, s1, c1, valid);
                else
                    intra_unit<true>(src, p.rtag + i0, n, q.x, q.y, nq, wrec, wtag, p, tab, s0, c0, s1, c1, valid);
                v = valid ? ((double)s0 - (double)c0) + ((double)s1 - (double)c1) : 0.0;
            } else {
                // the moved residue's pairs, new minus current
                float s0 = 0.f, c0 = 0.f;
                moved_item(src, p.pos + (size_t)r * p.stride + i0, p.rtag + i0, n, mv.y, l2, p, tab, s0, c0);
                v = (double)s0 - (double)c0;
            }
            __builtin_amdgcn_s_setprio(2);
            v = wave_total(v);
            if (lane == 0) unsafeAtomicAdd(p.rowbuf + (size_t)r * p.n_mol + m, v);
            continue;
        }
        const int ib = local / JC, jc = local % JC;
        const int m = p.prop_mol[r];
        const int n_m = p.mol_off[m + 1] - p.mol_off[m];
        const int jbase = jc * kJChunk;
        __syncthreads();  // previous item finished reading jrec
        {
            const int j = jbase + threadIdx.x;
            if (j < p.n_res) jrec[threadIdx.x] = p.pos[(size_t)r * p.stride + j];
        }
        const int il = ib * kIBlock + lane;
        const bool valid = il < n_m;
        const float4 pn = p.prop_pos[(size_t)r * p.max_n + (valid ? il : 0)];
        const uint32_t lane_off = (__float_as_uint(pn.w) >> kRowShift & 0x1Fu) << kColShift;
        const float4 ci = p.prop_center[r];
        const float ri = p.prop_radius[r];
        __syncthreads();
        int js = jbase + wave * kWave;
        const int je = min(p.n_res, jbase + (wave + 1) * kWave);
        __builtin_amdgcn_s_setprio(0);
        while (js < je) {
            const int J = p.mol_of[js];
            const int se = min(je, p.mol_off[J + 1]);
            if (J != m) {
                float s0 = 0.f, c0 = 0.f, s1 = 0.f, c1 = 0.f;
                segment<false>(jrec, js - jbase, se - jbase, jbase, pn, lane_off, tab, p, ci, ri,
                               p.center[(size_t)r * p.n_mol + J], p.radius[(size_t)r * p.n_mol + J],
                               m >= p.crowd_mol || J >= p.crowd_mol, s0, c0, s1, c1);
                double v = valid ? (double)s0 - (double)c0 : 0.0;
                v = wave_total(v);
                if (lane == 0) unsafeAtomicAdd(p.rowbuf + (size_t)r * p.n_mol + J, v);
            }
            js = se;
        }
        __builtin_amdgcn_s_setprio(2);
    }
}

// ---------------------------------------------------------------------------
// Delta-E row, j-lane layout (systems without flexible molecules).  For
// replica r with molecule m (n_m residues) moved, the j axis is every
// residue outside m in device order (n_j = n_res - n_m), cut into blocks
// of 64 j residues, one per lane.  One unit of work is an (i, j-block)
// step: all 64 lanes evaluate their j residue against residue i of m's
// proposal, read as a wave-uniform record through the scalar cache.  The
// units of all replicas are laid end to end (replica, j-block, i) and every
// wave takes an equal contiguous share, so the grid is balanced to within
// one unit and waves never wait for one another.  Each lane loads its
// table entry per i residue (residue types change at almost every residue);
// uncharged i residues skip the Debye-Hueckel term.  Per lane the terms are
// summed in groups of kRowGroup i residues into a Kahan-compensated FP32
// accumulator; at the end of a j-block visit the lanes of each molecule J
// are reduced in FP64 and added atomically into the row.
// ---------------------------------------------------------------------------

// Image handling of the row kernel, wave-uniform per j-block visit.  Each
// lane's j residue is moved to its image nearest the proposal's centre cm:
// e = cm - x_j with |e| <= L/2 on every axis, and every pair difference
// d = x_i - x_j = (x_i - cm) + e lies within rm of e.  On an axis where
// |e| < L/2 - rm for every lane, d is the minimum image.  Elsewhere (IMG
// bit set) a lane with |e| >= L/2 - rm ("straddling": its pairs may cross
// L/2) stores x_j + sign(e) L/2 and c = L/2, the others x_j and c = 0; the
// pair loop forms t = x_i - (stored x_j) and c - |t|, whose square is the
// square of the minimum image in one operation: for a straddling lane,
// t = d - sign(e) L/2 with sign(e) d in (-L/2, L) (rm < L/2), where
// L/2 - |t| is d on the near side and L - |d| on the far side; for the
// others c - |t| = -|d|.  The straddling lanes' pairs lie at least
// L/2 - 2 rm apart, where the rounding of the half-box shift is
// immaterial.  kImgFull: rm too large for these bounds; unshifted
// coordinates and the per-pair rint on every axis.
constexpr int kImgFull = 8;

template <int IMG>
__device__ __forceinline__ float image_axis(float d, float c, float box, float inv_box, int axis_bit)
{
    if (IMG == kImgFull) return fmaf(-box, rintf(d * inv_box), d);
    if (IMG & axis_bit) return c - fabsf(d);
    return d;
}

// The JPL j residues a lane holds (64 apart) with their table columns and
// the image constants of straddling lanes (see image_axis).
template <int JPL>
struct LaneJ {
    float x[JPL], y[JPL], z[JPL];
    float cx[JPL], cy[JPL], cz[JPL];
    uint32_t col[JPL];
};

// Squared minimum-image distance of a row-kernel pair: wave-uniform i
// record (xi, yi, zi), the lane's j residue h.
template <int IMG, int JPL>
__device__ __forceinline__ float row_r2(float xi, float yi, float zi, const LaneJ<JPL> &lj, int h, float box,
                                        float inv_box)
{
    const float dx = image_axis<IMG>(xi - lj.x[h], lj.cx[h], box, inv_box, 1);
    const float dy = image_axis<IMG>(yi - lj.y[h], lj.cy[h], box, inv_box, 2);
    const float dz = image_axis<IMG>(zi - lj.z[h], lj.cz[h], box, inv_box, 4);
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

// g plus one pair term of the row kernel at squared distance r2; the
// Debye-Hueckel term for charged i residues only.  FAR: the visit's pairs
// all lie beyond every switch distance, where the term is the attractive
// form.
template <bool CHARGED, bool FAR>
__device__ __forceinline__ float row_term(float r2, const RowCoef &cf, float g)
{
    float rinv = 0.0f, rinv2;
    if (CHARGED) {
        rinv = __builtin_amdgcn_rsqf(r2);
        rinv2 = rinv * rinv;
    } else {
        rinv2 = __builtin_amdgcn_rcpf(r2);
    }
    const float u = rinv2 * rinv2 * rinv2;
    const float w = fmaf(cf.B, u, -cf.C);
    float e;
    if (FAR) {
        e = fmaf(w, u, g);
    } else {
        const float att = w * u;
        e = g + (fmaf(cf.C, u, cf.K) < 0.0f ? cf.K - att : att);
    }
    if (CHARGED) e = fmaf(cf.qq * __builtin_amdgcn_exp2f(r2 * rinv * kDHExp2), rinv, e);
    return e;
}

// The i residues [ia, ib) of the proposal against the lane's j residues:
// terms summed plainly per j over kRowFlush groups of kRowGroup consecutive
// i residues (across type changes), the plain sums added into
// Kahan-compensated accumulators.  Each i residue's word (the record's w
// lane) gives its type, hence the lane's table entry (for far visits its
// first two or three words), and whether it is charged (Debye-Hueckel
// terms only then).  Records are read kRowGroup at a time through the
// scalar cache; the proposal buffers end with 8 spare records, so that a
// short group may read past a molecule's last residue (those terms are
// skipped).  Groups of 8 measured 2.7 % faster than groups of 4 on a
// 40-chain system, and one compensated add per two groups 0.7 % faster
// than one per group.
constexpr int kRowGroup = 8;
constexpr int kRowFlush = 2;
// (a group starting at the last residue reads kRowGroup - 1 records past it)
static_assert(kRowGroup - 1 <= 8, "the proposal buffers end with 8 spare records");
template <int IMG, bool FAR, int JPL>
__device__ __forceinline__ void row_residue(const PairParams &p, const f4_t &ri, const LaneJ<JPL> &lj,
                                            const char *__restrict__ tab, uint32_t &cur, RowCoef (&cf)[JPL],
                                            float (&gu)[JPL], float (&gc)[JPL])
{
    const uint32_t code = __float_as_uint(ri.w);
    // the lane's table entries change with the residue type only (each
    // molecule's residues are sorted by charge and type: runs of about 8
    // residues); the loads are issued before the distances
    const bool fresh = (code & kRowMask) != cur;
    if (fresh) {
        cur = code & kRowMask;
#pragma unroll
        for (int h = 0; h < JPL; ++h) cf[h] = *reinterpret_cast<const RowCoef *>(tab + (cur | lj.col[h]));
    }
    float r2[JPL];
#pragma unroll
    for (int h = 0; h < JPL; ++h) r2[h] = row_r2<IMG, JPL>(ri.x, ri.y, ri.z, lj, h, p.box, p.inv_box);
    // (each branch adds into an accumulator of its own, which it updates
    // in place: one shared accumulator costs register copies at the join)
    if (code & kChargedBit) {
#pragma unroll
        for (int h = 0; h < JPL; ++h) gc[h] = row_term<true, FAR>(r2[h], cf[h], gc[h]);
    } else {
#pragma unroll
        for (int h = 0; h < JPL; ++h) gu[h] = row_term<false, FAR>(r2[h], cf[h], gu[h]);
    }
}

template <int IMG, bool FAR, int JPL>
__device__ __forceinline__ void row_runs(const PairParams &p, const float4 *__restrict__ irec, int ia, int ib,
                                         const LaneJ<JPL> &lj, const char *__restrict__ tab, float (&s)[JPL],
                                         float (&c)[JPL])
{
    uint32_t cur = ~0u;  // no residue type yet
    RowCoef cf[JPL];
    float gu[JPL], gc[JPL];
#pragma unroll
    for (int h = 0; h < JPL; ++h) gu[h] = gc[h] = 0.0f;
    int k = 0;
    for (int i = ia; i < ib; i += kRowGroup) {
        f4_t r[kRowGroup];
#pragma unroll
        for (int u = 0; u < kRowGroup; ++u) r[u] = scalar_rec(irec, i + u);
        const int n = ib - i;
        row_residue<IMG, FAR, JPL>(p, r[0], lj, tab, cur, cf, gu, gc);
#pragma unroll
        for (int u = 1; u < kRowGroup; ++u)
            if (u < n) row_residue<IMG, FAR, JPL>(p, r[u], lj, tab, cur, cf, gu, gc);
        // the plain sums enter the compensated ones every kRowFlush groups
        // and after the last
        if (++k == kRowFlush || n <= kRowGroup) {
#pragma unroll
            for (int h = 0; h < JPL; ++h) {
                kahan_add(s[h], c[h], gu[h] + gc[h]);
                gu[h] = gc[h] = 0.0f;
            }
            k = 0;
        }
    }
}

// Proposal records of window slot s of replica r (slot 0: the pending
// proposal).
__device__ __forceinline__ float4 *slot_pos(const PairParams &p, int r, int s)
{
    return s == 0 ? p.prop_pos + (size_t)r * p.max_n
                  : p.wprop_pos + ((size_t)r * ((1 << p.win_log2) - 1) + (s - 1)) * p.max_n;
}
__device__ __forceinline__ float4 *slot_rel(const PairParams &p, int r, int s)
{
    return s == 0 ? p.prop_rel + (size_t)r * p.max_n
                  : p.wprop_rel + ((size_t)r * ((1 << p.win_log2) - 1) + (s - 1)) * p.max_n;
}

// Proposal headers the row kernel keeps in LDS (at most this many).
constexpr int kHdrLds = 128;

// CROWD: the system has crowders under repulsive crowding (a separate
// instantiation, so that systems without them keep the registers of the
// plain kernel).  WIN: speculative windows; the kernel's "proposals" are
// the R W window slots v = r W + s, each evaluated against replica r's
// stored residues outside its molecule and against the proposals of the
// slots before it (J = M + q in the window row).
// (the crowder instantiations hold the crowder pass's state as well: below
// 8 waves per SIMD rather than spill)
template <int JPL, bool CROWD, bool WIN>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(JPL > 2 || CROWD ? 4 : 8, 8))) void mc_row_kernel(PairParams p)
{
    static_assert(!(CROWD && WIN), "windows run without repulsive crowding");
    // Wave priority: the SIMDs issue from the oldest wave first, so a wave
    // still in its set-up (a chain of dependent loads and a few hundred
    // instructions) would wait behind the pair loops of older waves and
    // start its own loops last.  Set-up, visit headers and reductions run
    // at high priority, the pair loops at normal priority.
    __builtin_amdgcn_s_setprio(2);
    __shared__ __attribute__((aligned(16))) char tab[kTableBytes];
    // dynamic LDS (row_lds_bytes): the proposals' unit offsets, then for at
    // most kHdrLds proposals their headers and centres
    extern __shared__ __attribute__((aligned(16))) long long row_lds[];
    const int lw = WIN ? p.win_log2 : 0;
    const int R = p.rep_count;
    const int V = R << lw;  // proposals: replicas, or window slots
    const bool hdr_lds = V <= kHdrLds;
    long long *unit_off = row_lds;
    int4 *s_hdr = reinterpret_cast<int4 *>(row_lds + ((V + 2) & ~1));
    float4 *s_cen = reinterpret_cast<float4 *>(s_hdr + (hdr_lds ? V : 0));
    const int4 *g_hdr = WIN ? p.whdr : p.prop_hdr;
    const float4 *g_cen = WIN ? p.wcen : p.prop_center;
    // every set-up load is issued before the LDS stores: the table (entry
    // e = tj * 20 + ti, i = the moved residue, goes to row ti, column tj),
    // and the first 64 proposals' headers and centres
    static_assert(CGPPD_AA_COUNT * CGPPD_AA_COUNT <= 2 * kWG, "two table entries per thread");
    // (unconditional, so that none waits for another; only wave 0 uses
    // the headers)
    const int e0 = threadIdx.x, e1 = threadIdx.x + kWG;
    const int rr = min((int)threadIdx.x & (kWave - 1), V - 1);
    const RowCoef t0 = p.row_table[e0];
    const RowCoef t1 = p.row_table[min(e1, CGPPD_AA_COUNT * CGPPD_AA_COUNT - 1)];
    const int4 hd0 = g_hdr[rr];
    const float4 cen0 = g_cen[rr];
    const int wl0 = WIN ? p.wlen[rr >> lw] : 0;
    constexpr int kEntries = CGPPD_AA_COUNT * CGPPD_AA_COUNT;
    const int e1c = min(e1, kEntries - 1);  // (threads past the table rewrite its last entry)
    *reinterpret_cast<RowCoef *>(tab + (((e0 % CGPPD_AA_COUNT) << kRowShift) | ((e0 / CGPPD_AA_COUNT) << kColShift))) = t0;
    *reinterpret_cast<RowCoef *>(tab + (((e1c % CGPPD_AA_COUNT) << kRowShift) | ((e1c / CGPPD_AA_COUNT) << kColShift))) = t1;
    // unit offsets of the proposals: an inclusive scan, 64 proposals at a
    // time, from the proposal headers (the first 64 loaded above).  Every
    // wave scans (the loads then have an unconditional use and are not
    // deferred); wave 0 stores.
    {
        const int lane0 = threadIdx.x & (kWave - 1);
        const bool w0 = threadIdx.x < kWave;
        // a proposal's j residues: the other molecules' residues, plus,
        // for a window slot, the residues of the slots before it (cen.w)
        auto units = [&](int v, int4 hd, float4 cen, int wl) -> long long {
            if (v >= V) return 0;
            if (WIN) {
                if ((v & ((1 << lw) - 1)) >= wl) return 0;
            } else if (p.da && !p.s1[v]) {
                return 0;
            }
            const int nj = p.n_res - hd.z + (WIN ? __float_as_int(cen.w) : 0);
            return (long long)((nj + JPL * kWave - 1) / (JPL * kWave)) * hd.z;
        };
        long long c = wave_scan_i64(units(lane0, hd0, cen0, wl0));
        if (w0 && lane0 < V) {
            unit_off[lane0 + 1] = c;
            if (hdr_lds) {
                s_hdr[lane0] = hd0;
                s_cen[lane0] = cen0;
            }
        }
        long long carry = readlane_b64(c, kWave - 1);
        for (int base = kWave; base < V; base += kWave) {
            const int v = base + lane0, vc = min(v, V - 1);
            const int4 hd = g_hdr[vc];
            const float4 cen = g_cen[vc];
            c = wave_scan_i64(units(v, hd, cen, WIN ? p.wlen[vc >> lw] : 0));
            if (w0 && v < V) {
                unit_off[v + 1] = carry + c;
                if (hdr_lds) {
                    s_hdr[v] = hd;
                    s_cen[v] = cen;
                }
            }
            carry += readlane_b64(c, kWave - 1);
        }
        if (threadIdx.x == 0) unit_off[0] = 0;
    }
    __syncthreads();
    const long long total = unit_off[V];
    const int G = gridDim.x, b = blockIdx.x;
    const int lb = (p.xcd_map && (G & 7) == 0) ? (b & 7) * (G >> 3) + (b >> 3) : b;
    const int lane = threadIdx.x & (kWave - 1);
    const int W = G * kWavesPerWG;
    const int wv = lb * kWavesPerWG + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // share boundaries floor(total w / W) in FP64 (exact while total W
    // < 2^53; every wave evaluates the same expression for a shared
    // boundary, so the shares tile [0, total) exactly)
    const double per = (double)total / (double)W;
    long long P = wv == 0 ? 0 : (long long)floor((double)wv * per);
    const long long P1 = wv + 1 == W ? total : (long long)floor((double)(wv + 1) * per);
    if (P >= P1) return;
    // proposal holding unit P: the number of proposals that end at or
    // before P (unit_off is non-decreasing and P < unit_off[V]), 64 at a time
    int v = 0;
    for (int base = 0; base < V; base += kWave) {
        const int k = base + lane;
        const unsigned long long before = __ballot(k < V && unit_off[k + 1] <= P);
        v = base + __popcll(before);
        if (before != ~0ull) break;
    }
    v = __builtin_amdgcn_readfirstlane(v);
    const int M = p.n_mol;
    // position of unit P inside its proposal: j-block jb, residue ia of m
    // (one division for the first visit, then the visits advance)
    int jb = -1, ia = 0;
    while (P < P1) {
        if (jb < 0 || P >= unit_off[v + 1]) {
            while (P >= unit_off[v + 1]) ++v;
            jb = -1;
        }
        // replica r and window slot s of proposal v
        const int r = v >> lw, slot = v & ((1 << lw) - 1);
        // the proposal: molecule, first residue, size, bounding radius, centre
        int4 hd;
        float4 cm;
        if (hdr_lds) {
            hd = s_hdr[v];
            cm = s_cen[v];
        } else {
            hd = g_hdr[v];
            cm = g_cen[v];
        }
        const int i0 = __builtin_amdgcn_readfirstlane(hd.y);
        const int n_m = __builtin_amdgcn_readfirstlane(hd.z);
        const float rm = __int_as_float(__builtin_amdgcn_readfirstlane(hd.w));
        const int ext = WIN ? __builtin_amdgcn_readfirstlane(__float_as_int(cm.w)) : 0;
        cm.x = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cm.x)));
        cm.y = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cm.y)));
        cm.z = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(cm.z)));
        if (jb < 0) {
            // units per proposal fit in 32 bits (set_replicas bounds them)
            const unsigned local = (unsigned)(P - unit_off[v]);
            jb = __builtin_amdgcn_readfirstlane((int)(local / (unsigned)n_m));
            ia = __builtin_amdgcn_readfirstlane((int)(local % (unsigned)n_m));
        }
        const int ib = __builtin_amdgcn_readfirstlane((int)min((long long)n_m, (long long)ia + (P1 - P)));
        const float4 *irec = WIN ? slot_pos(p, r, slot) : p.prop_pos + (size_t)r * p.max_n;
        // this lane's j residues (jb * 64 JPL + 64 h + lane, h < JPL): the
        // stored residues outside m, then (window slots) the earlier
        // slots' proposed residues
        const int nmain = p.n_res - n_m;
        const int nj = nmain + ext;
        LaneJ<JPL> lj;
        bool valid[JPL];
        int J[JPL];
        float4 rj[JPL];
        // (every load issued before any is used; lanes past the last j
        // residue read the last residue and are masked)
        if (!WIN || (jb + 1) * (JPL * kWave) <= nmain) {
#pragma unroll
            for (int h = 0; h < JPL; ++h) {
                const int jp = jb * (JPL * kWave) + h * kWave + lane;
                valid[h] = jp < nj;
                const int j = min(jp + (jp >= i0 ? n_m : 0), p.n_res - 1);
                rj[h] = p.pos[(size_t)r * p.stride + j];
                J[h] = p.mol_of[j];
            }
        } else {
            // a j-block reaching into the earlier slots' proposals: slot q
            // holds window j indices [nmain + cum_q, nmain + cum_q + n_q)
#pragma unroll
            for (int h = 0; h < JPL; ++h) {
                const int jp = jb * (JPL * kWave) + h * kWave + lane;
                valid[h] = jp < nj;
                if (jp < nmain) {
                    const int j = jp + (jp >= i0 ? n_m : 0);
                    rj[h] = p.pos[(size_t)r * p.stride + j];
                    J[h] = p.mol_of[j];
                } else if (jp >= nj) {
                    rj[h] = p.pos[(size_t)r * p.stride + p.n_res - 1];
                    J[h] = 0;
                } else {
                    // 0 <= k < ext: slot >= 1 and slot q < slot holds k
                    const int k = jp - nmain;
                    int q = 0, cum = 0;
                    for (; q < slot - 1; ++q) {
                        const int nq = g_hdr[(r << lw) + q].z;
                        if (k < cum + nq) break;
                        cum += nq;
                    }
                    rj[h] = slot_pos(p, r, q)[k - cum];
                    J[h] = M + q;
                }
            }
        }
#pragma unroll
        for (int h = 0; h < JPL; ++h) {
            if (!valid[h]) {
                rj[h] = make_float4(0.f, 0.f, 0.f, 0.f);
                J[h] = 0x7fffffff;
            }
            lj.col[h] = lane_col(rj[h]);
        }
        // repulsive crowding: a visit pairing a crowder with anything (the
        // moved molecule is one, or some lane's molecule is) adds the
        // crowder pass below, whose sums replace the crowder lanes' own
        const bool m_crowd = CROWD && __builtin_amdgcn_readfirstlane(hd.x) >= p.crowd_mol;
        bool cl[JPL];
        bool any_c = false;
#pragma unroll
        for (int h = 0; h < JPL; ++h) {
            cl[h] = CROWD && valid[h] && (m_crowd || J[h] >= p.crowd_mol);
            any_c |= cl[h];
        }
        const bool crowd_visit = CROWD && __any(any_c);
        // image handling of this visit (see image_axis): each j residue is
        // moved to its image nearest the proposal's centre cm; every moved
        // residue i lies within rm of cm, so on an axis where
        // |cm - x_j| < L/2 - rm (with margin) for every lane the plain
        // difference is the minimum image, and elsewhere
        // min(|dx|, L - |dx|) is (|dx| < L needs rm < L/2, with margin)
        //
        // far visit: every j residue lies beyond rm + lj_reach of cm, so
        // every pair is past its switch distance (the pair loops then skip
        // the repulsive branch)
        int img = kImgFull;
        bool far = false;
        // per lane and axis: the image shift, the straddling half-box
        // shift and the image constant c
        float sh[JPL][3], hs[JPL][3], hc[JPL][3];
        const float reach = rm * 1.0001f + 0.01f;
        if (p.fast_image && reach < p.image_reach) {
            const float lim = p.half_box - reach, half = p.half_box;
            const float fr = reach + p.lj_reach;
            int need = 0;
            bool near = false;
#pragma unroll
            for (int h = 0; h < JPL; ++h) {
#pragma unroll
                for (int a = 0; a < 3; ++a) sh[h][a] = hs[h][a] = hc[h][a] = 0.0f;
                if (!valid[h]) continue;
                const float o[3] = {cm.x - rj[h].x, cm.y - rj[h].y, cm.z - rj[h].z};
                float e[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float k = rintf(o[a] * p.inv_box);
                    e[a] = fmaf(-p.box, k, o[a]);
                    sh[h][a] = p.box * k;
                    const bool st = !(fabsf(e[a]) < lim);
                    need |= st ? 1 << a : 0;
                    hs[h][a] = st ? copysignf(half, e[a]) : 0.0f;
                    hc[h][a] = st ? half : 0.0f;
                }
                near |= fmaf(e[0], e[0], fmaf(e[1], e[1], e[2] * e[2])) <= fr * fr;
            }
            img = (__any(need & 1) ? 1 : 0) | (__any(need & 2) ? 2 : 0) | (__any(need & 4) ? 4 : 0);
            far = !__any(near);
        }
        const bool shift = img != kImgFull;
#pragma unroll
        for (int h = 0; h < JPL; ++h) {
            lj.x[h] = shift ? rj[h].x + sh[h][0] + hs[h][0] : rj[h].x;
            lj.y[h] = shift ? rj[h].y + sh[h][1] + hs[h][1] : rj[h].y;
            lj.z[h] = shift ? rj[h].z + sh[h][2] + hs[h][2] : rj[h].z;
            lj.cx[h] = shift ? hc[h][0] : 0.0f;
            lj.cy[h] = shift ? hc[h][1] : 0.0f;
            lj.cz[h] = shift ? hc[h][2] : 0.0f;
        }
        float sa[JPL], ca[JPL];
#pragma unroll
        for (int h = 0; h < JPL; ++h) sa[h] = ca[h] = 0.0f;
        __builtin_amdgcn_s_setprio(0);
        // far visits: cases 0-7 (the image axes); near visits: 16 + axes
        switch (m_crowd ? -1 : img == kImgFull ? kImgFull : far ? img : 16 + img) {
        case -1: break;  // every pair of the visit is a crowder pair
        case 0: row_runs<0, true, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 1: row_runs<1, true, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 2: row_runs<2, true, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 3: row_runs<3, true, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 4: row_runs<4, true, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 5: row_runs<5, true, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 6: row_runs<6, true, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 7: row_runs<7, true, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 16: row_runs<0, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 17: row_runs<1, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 18: row_runs<2, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 19: row_runs<3, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 20: row_runs<4, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 21: row_runs<5, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 22: row_runs<6, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        case 23: row_runs<7, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        default: row_runs<kImgFull, false, JPL>(p, irec, ia, ib, lj, tab, sa, ca); break;
        }
        if (crowd_visit) {
            // crowder pairs: the stored j coordinates (reloaded: the
            // common visit keeps no copy), per-pair rint image, kcal/mol
            float4 q[JPL];
#pragma unroll
            for (int h = 0; h < JPL; ++h) {
                const int jp = jb * (JPL * kWave) + h * kWave + lane;
                q[h] = p.pos[(size_t)r * p.stride + min(jp + (jp >= i0 ? n_m : 0), p.n_res - 1)];
            }
            float sc[JPL], cc[JPL];
#pragma unroll
            for (int h = 0; h < JPL; ++h) sc[h] = cc[h] = 0.0f;
            for (int i = ia; i < ib; ++i) {
                const f4_t ri = scalar_rec(irec, i);
#pragma unroll
                for (int h = 0; h < JPL; ++h) {
                    const float e = crowd_term(min_image(ri.x - q[h].x, p.box, p.inv_box),
                                               min_image(ri.y - q[h].y, p.box, p.inv_box),
                                               min_image(ri.z - q[h].z, p.box, p.inv_box));
                    kahan_add(sc[h], cc[h], e * p.rt_to_kcal);
                }
            }
#pragma unroll
            for (int h = 0; h < JPL; ++h)
                if (cl[h]) {
                    sa[h] = sc[h];
                    ca[h] = cc[h];
                }
        }
        __builtin_amdgcn_s_setprio(2);
        // FP64 reduction per molecule J: across the lanes J[h] is
        // non-decreasing for each h, and every J[h] <= every J[h + 1]
        double vs[JPL];
#pragma unroll
        for (int h = 0; h < JPL; ++h) vs[h] = valid[h] ? (double)sa[h] - (double)ca[h] : 0.0;
        // (lane 0 holds the block's first j residue: when it has none, no
        // lane has)
        int Jc = __builtin_amdgcn_readfirstlane(J[0]);
        unsigned long long *const fix_row = WIN ? p.wrowfix + (size_t)v * (M + (1 << lw)) : p.rowfix + (size_t)r * M;
        double *const dbl_row = WIN ? p.wrowbuf + (size_t)v * (M + (1 << lw)) : p.rowbuf + (size_t)r * M;
        while (Jc != 0x7fffffff) {
            double part = 0.0;
#pragma unroll
            for (int h = 0; h < JPL; ++h) part += J[h] == Jc ? vs[h] : 0.0;
            const double t = wave_total(part);
            if (lane == 0) {
                if (fabs(t) < kFixMax)
                    atomicAdd(fix_row + Jc, (unsigned long long)__double2ll_rn(t * kFixScale));
                else
                    unsafeAtomicAdd(dbl_row + Jc, t);
            }
            int next = -1;
#pragma unroll
            for (int h = 0; h < JPL; ++h) {
                if (next >= 0) break;
                const unsigned long long rest = __ballot(valid[h] && J[h] > Jc);
                if (rest) next = __builtin_amdgcn_readlane(J[h], __builtin_ctzll(rest));
            }
            Jc = next < 0 ? 0x7fffffff : next;
        }
        P += ib - ia;
        // the next visit starts at the next j-block of this proposal, or at
        // the first unit of a later one
        if (ib == n_m) {
            ++jb;
            ia = 0;
        } else {
            ia = ib;
        }
    }
}

// Items of the i-lane pair kernel for a moved molecule m of n residues:
// (i-block, j-chunk) items, and for a flexible molecule its intra-molecular
// items: for kMovedAll every (i-block, j-chunk) item, for kMovedOne one item
// per j-chunk.
__device__ __forceinline__ int ilane_items(const PairParams &p, int m, int n, int kind)
{
    const int nib = (n + kIBlock - 1) / kIBlock;
    int items = p.n_flex ? ilane_inter_items(p, nib, n) : nib * p.jc_count;
    if (p.internal && p.n_flex && p.mol_flex[m])
        items += kind == kMovedOne ? (n + kJChunk - 1) / kJChunk : (intra_units(n) + kWavesPerWG - 1) / kWavesPerWG;
    return items;
}

// Thread 0, after a proposal of molecule m (n residues) of replica r is
// laid out: its i-lane items, the internal-energy record, the proposed
// bounding radius, and the diagonal row entry the pair kernel adds the
// intra-molecular pairs to: the bonded energy of the proposal (kMovedAll),
// or the current internal energy plus the bonded change (kMovedOne); bonded
// is that value in kcal/mol.
__device__ __forceinline__ void record_proposal(const PairParams &p, int r, int m, int i0, int n, int kind,
                                                int moved, float radius_new, double bonded)
{
    const bool flex = p.internal && p.n_flex && p.mol_flex[m];
    if (!flex) kind = kMovedNone;
    p.prop_hdr[r] = make_int4(m, i0, n, __float_as_int(radius_new));
    p.prop_items[r] = ilane_items(p, m, n, kind);
    p.prop_moved[r] = make_int4(kind, moved, 0, 0);
    p.prop_radius[r] = radius_new;
    if (flex) {
        const size_t d = (size_t)r * p.n_mol + m;
        p.rowbuf[d] = (kind == kMovedOne ? p.emat[(size_t)r * p.n_mol * p.n_mol + (size_t)m * p.n_mol + m] : 0.0) + bonded;
    }
}

// Bonded energy (kcal/mol) of flexible molecule m's proposal of replica r,
// or, when only one residue moved (molecule-local index moved >= 0), the
// proposal's minus that of the current coordinates (whole workgroup; every
// thread returns the sum).  The proposal is in place.
__device__ double proposal_bonded(const PairParams &p, int r, int m, int moved)
{
    __shared__ double sb[kWavesPerWG];
    const int i0 = p.mol_off[m];
    double tn[5];
    bonded_terms<true>(p.prop_pos + (size_t)r * p.max_n, i0, m, p, nullptr, tn,
                       moved >= 0 ? p.pos + (size_t)r * p.stride + i0 : nullptr, moved);
    double v = bonded_kcal(tn);
    v = wave_total(v);
    if ((threadIdx.x & (kWave - 1)) == 0) sb[threadIdx.x / kWave] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kWavesPerWG; ++w) t += sb[w];
    __syncthreads();  // sb may be reused by the next call
    return t;
}

// The topology side of a flexible move (kind 2 flex, 3 local) of molecule
// m, fixed by the molecule and the move's random words alone: the flex
// move's pivot and branch, or the local move's crankshaft triple or
// translated residue (device indices; -1 where the molecule has none).
// It takes a chain of dependent index loads, which mc_finish_kernel
// resolves for the next step while the current step is decided.
struct FlexPlan {
    uint32_t x[4];  // the move's Philox words
    int pivot;      // flex: the flexible residue the branch turns about
    int b0, nb;     // flex: the branch residues br[b0, b0 + nb)
    int4 cr;        // crankshaft: residues i - 1, i, i + 1
    int li;         // local translation: the residue
};

__device__ __forceinline__ void plan_flexible(const PairParams &p, int r, int64_t step, int m, int kind, FlexPlan &fp)
{
    philox((uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)(r + p.rep_rng), kTagMC | 2u, (uint32_t)p.seed,
           (uint32_t)(p.seed >> 32), fp.x);
    fp.pivot = -1;
    fp.b0 = fp.nb = 0;
    fp.cr = make_int4(-1, -1, -1, -1);
    fp.li = -1;
    if (kind == 2) {
        // flex(axis, rotate step, ri, neighbour): src/Molecule.cpp:331-413
        const int nf = p.flexr_off[m + 1] - p.flexr_off[m];
        if (nf > 0) {
            const int e = p.flexr_off[m] + mulhi_n(fp.x[0], nf);
            const int deg = p.nb_off[e + 1] - p.nb_off[e];
            const int k = p.nb_off[e] + mulhi_n(fp.x[1], deg);
            fp.pivot = p.flexr[e];
            fp.b0 = p.br_off[k];
            fp.nb = p.br_off[k + 1] - fp.b0;
        }
    } else if (fp.x[2] < 0x80000000u) {
        // crankshaft: src/Molecule.cpp:303-329, 415-448
        const int nc = p.crank_off[m + 1] - p.crank_off[m];
        if (nc > 0) fp.cr = p.crank[p.crank_off[m] + mulhi_n(fp.x[0], nc)];
    } else {
        // local translation: src/Molecule.cpp:422-430
        const int nl = p.local_off[m + 1] - p.local_off[m];
        if (nl > 0) fp.li = p.local[p.local_off[m] + mulhi_n(fp.x[0], nl)];
    }
}

// A flexible move of molecule m (kind 2 flex, 3 local) of replica r, laid
// out as the proposal (whole workgroup; src/Molecule.cpp:274-448).  The
// current coordinates are copied, the residues that change are moved (a
// flex branch in parallel, one residue per thread, each displacement kept
// in LDS), thread 0 sums the displacements in the caller's order,
// then every residue's relative position takes the centre shift and the
// molecule is wrapped (src/Molecule.cpp:242-252).  Bit-identical to the CPU
// restatement: no FMA contraction, correctly rounded division and square
// root, the same summation order.
__device__ void propose_flexible(const PairParams &p, int r, int64_t step, const Proposal &pr, int kind,
                                 const FlexPlan *plan, const int *plan_br, bool copied)
{
#pragma clang fp contract(off)
    __shared__ float sh_vf[3], sh_nc[3];
    __shared__ int sh_moved;
    const int m = pr.mol;
    const int i0 = p.mol_off[m], n = p.mol_off[m + 1] - i0;
    const float4 *pos = p.pos + (size_t)r * p.stride + i0;
    const float4 *rel = p.rel + (size_t)r * p.stride + i0;
    float4 *pp = p.prop_pos + (size_t)r * p.max_n;
    float4 *pq = p.prop_rel + (size_t)r * p.max_n;
    __shared__ int sh_mone, sh_wrap;
    __shared__ float sh_r2[kWavesPerWG];
    // the molecule's current centre and radius, for thread 0 below
    float4 c_cur = make_float4(0.f, 0.f, 0.f, 0.f);
    float rad_cur = 0.0f;
    if (threadIdx.x == 0) {
        c_cur = p.center[(size_t)r * p.n_mol + m];
        rad_cur = p.radius[(size_t)r * p.n_mol + m];
    }
    // the proposal starts as a copy of the current coordinates (already in
    // place when copied)
    if (!copied)
        for (int k = threadIdx.x; k < n; k += kWG) {
            pp[k] = pos[k];
            pq[k] = rel[k];
        }
    if (threadIdx.x == 0) sh_mone = -1;
    // the move's topology (every thread resolves the same plan unless the
    // caller did)
    FlexPlan fp;
    if (plan)
        fp = *plan;
    else
        plan_flexible(p, r, step, m, kind, fp);
    __syncthreads();
    const uint32_t x[4] = {fp.x[0], fp.x[1], fp.x[2], fp.x[3]};
    const double rcos = p.rcos[r], rsin = p.rsin[r];
    float acc[3] = {0.f, 0.f, 0.f};
    int moved = 0;
    if (kind == 2) {
        // flex(axis, rotate step, ri, neighbour): src/Molecule.cpp:331-413
        if (fp.pivot >= 0) {
            const double ax = pr.axis[0], ay = pr.axis[1], az = pr.axis[2];
            const double mag = sqrt(ax * ax + ay * ay + az * az);
            const double q[4] = {rcos, rsin * (ax / mag), rsin * (ay / mag), rsin * (az / mag)};
            const float4 c0 = pos[fp.pivot - i0];
            const int b0 = fp.b0, nb = fp.nb;
            // chunks of kWG residues: each thread moves one and keeps its
            // displacement in LDS, then thread 0 adds the chunk in order
            __shared__ float3 sdl[kWG];
            for (int base = 0; base < nb; base += kWG) {
                const int t = base + threadIdx.x;
                if (t < nb) {
                    const int li = (base == 0 && plan_br ? plan_br[threadIdx.x] : p.br[b0 + t]) - i0;
                    const float4 pi = pos[li];
                    const float3 rr = qrot(q, pi.x - c0.x, pi.y - c0.y, pi.z - c0.z);
                    const float dx = rr.x + c0.x - pi.x, dy = rr.y + c0.y - pi.y, dz = rr.z + c0.z - pi.z;
                    const float4 qi = rel[li];
                    pp[li] = make_float4(pi.x + dx, pi.y + dy, pi.z + dz, pi.w);
                    pq[li] = make_float4(qi.x + dx, qi.y + dy, qi.z + dz, qi.w);
                    sdl[threadIdx.x] = make_float3(dx, dy, dz);
                }
                __syncthreads();  // the chunk's displacements are in place
                if (threadIdx.x == 0) {
                    // eight LDS reads in flight per step, then the ordered adds
                    const int n = min(kWG, nb - base);
                    int k = 0;
                    for (; k + 8 <= n; k += 8) {
                        float3 d[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) d[u] = sdl[k + u];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            acc[0] = acc[0] + d[u].x;
                            acc[1] = acc[1] + d[u].y;
                            acc[2] = acc[2] + d[u].z;
                        }
                    }
                    for (; k < n; ++k) {
                        const float3 d = sdl[k];
                        acc[0] = acc[0] + d.x;
                        acc[1] = acc[1] + d.y;
                        acc[2] = acc[2] + d.z;
                    }
                }
                __syncthreads();  // the chunk was read before the next one is written
            }
            moved = 1;
        }
    } else if (threadIdx.x == 0) {
        // one local move, src/Molecule.cpp:415-448
        int li = -1;
        float v[3];
        if (x[2] < 0x80000000u) {
            // crankshaft(rotate step, flip, ri): src/Molecule.cpp:303-329
            if (fp.cr.x >= 0) {
                const int4 cr = fp.cr;
                const float4 a = pos[cr.x - i0], b = pos[cr.y - i0], c = pos[cr.z - i0];
                double ax = (double)(c.x - a.x), ay = (double)(c.y - a.y), az = (double)(c.z - a.z);
                const double mag = sqrt(ax * ax + ay * ay + az * az);
                ax /= mag;
                ay /= mag;
                az /= mag;
                const double sn = x[3] < 0x80000000u ? -rsin : rsin;
                double q[4] = {rcos, sn * ax, sn * ay, sn * az};
                const double qm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
                q[0] /= qm;
                q[1] /= qm;
                q[2] /= qm;
                q[3] /= qm;
                const float3 rr = qrot(q, b.x - a.x, b.y - a.y, b.z - a.z);
                v[0] = rr.x + a.x - b.x;
                v[1] = rr.y + a.y - b.y;
                v[2] = rr.z + a.z - b.z;
                li = cr.y - i0;
            }
        } else {
            // local translate by 0.1 x the translation step: src/Molecule.cpp:422-430
            if (fp.li >= 0) {
                li = fp.li - i0;
                float n0 = (float)pr.axis[0], n1 = (float)pr.axis[1], n2 = (float)pr.axis[2];
                const float mag = sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
                if (mag != 0.0f) {
                    n0 = n0 / mag;
                    n1 = n1 / mag;
                    n2 = n2 / mag;
                }
                const double sc = (double)0.1f * (double)p.tstep[r];
                v[0] = (float)((double)n0 * sc);
                v[1] = (float)((double)n1 * sc);
                v[2] = (float)((double)n2 * sc);
            }
        }
        if (li >= 0) {
            // translate(v, ri): src/Molecule.cpp:274-301
            const float4 pi = pos[li], qi = rel[li];
            pp[li] = make_float4(pi.x + v[0], pi.y + v[1], pi.z + v[2], pi.w);
            pq[li] = make_float4(qi.x + v[0], qi.y + v[1], qi.z + v[2], qi.w);
            acc[0] = v[0];
            acc[1] = v[1];
            acc[2] = v[2];
            moved = 1;
            sh_mone = li;
        }
    }
    if (threadIdx.x == 0) {
        const float4 c = c_cur;
        const float fn = (float)n;
        for (int k = 0; k < 3; ++k) sh_vf[k] = moved ? acc[k] / fn : 0.0f;
        sh_nc[0] = c.x + sh_vf[0];
        sh_nc[1] = c.y + sh_vf[1];
        sh_nc[2] = c.z + sh_vf[2];
        sh_moved = moved;
    }
    __syncthreads();
    const float vf0 = sh_vf[0], vf1 = sh_vf[1], vf2 = sh_vf[2];
    float c0 = sh_nc[0], c1 = sh_nc[1], c2 = sh_nc[2];
    float r2max = 0.0f;
    if (sh_moved) {
        const float w0 = fmodf(c0, p.box), w1 = fmodf(c1, p.box), w2 = fmodf(c2, p.box);
        const bool wrap = w0 != c0 || w1 != c1 || w2 != c2;
        for (int k = threadIdx.x; k < n; k += kWG) {
            float4 q = pq[k];
            q.x = q.x - vf0;
            q.y = q.y - vf1;
            q.z = q.z - vf2;
            pq[k] = q;
            if (wrap) pp[k] = make_float4(q.x + w0, q.y + w1, q.z + w2, pp[k].w);
            r2max = fmaxf(r2max, q.x * q.x + q.y * q.y + q.z * q.z);
        }
        if (wrap) {
            c0 = w0;
            c1 = w1;
            c2 = w2;
        }
        if (threadIdx.x == 0) sh_wrap = wrap;
    }
    // bounding radius of the proposed molecule (block maximum)
    r2max = wave_max_nonneg(r2max);
    if ((threadIdx.x & (kWave - 1)) == 0) sh_r2[threadIdx.x / kWave] = r2max;
    __syncthreads();
    const bool one = sh_moved && sh_mone >= 0 && !sh_wrap;
    const double bonded = proposal_bonded(p, r, m, one ? sh_mone : -1);
    if (threadIdx.x == 0) {
        float rad = rad_cur;
        if (sh_moved) {
            float q2 = 0.0f;
            for (int w = 0; w < kWavesPerWG; ++w) q2 = fmaxf(q2, sh_r2[w]);
            rad = sqrtf(q2) * 1.0001f;
        }
        p.prop_mol[r] = m;
        p.prop_nib[r] = (n + kIBlock - 1) / kIBlock;
        p.prop_n[r] = n;
        p.prop_center[r] = make_float4(c0, c1, c2, 0.0f);
        record_proposal(p, r, m, i0, n, one ? kMovedOne : kMovedAll, one ? sh_mone : 0, rad, bonded);
    }
}

// Lays out the proposal of MC step `step` for replica r (whole workgroup):
// molecule, new centre, and the proposed positions of its residues in
// device order (src/Molecule.cpp:146-266, src/Quaternion.cpp:12-112).
// Flexible molecules draw translate / rotate / flex / local with equal
// weights (rotate / flex / local when the molecule is alone,
// src/Molecule.cpp:451-466, src/Replica.cpp:309-334).
// plan, plan_br and copied: the caller resolved the flexible move's
// topology (plan_br: the first kWG branch residues of a flex move) and
// already copied the current coordinates into the proposal buffers.
__device__ void propose(const PairParams &p, int r, int64_t step, const FlexPlan *plan = nullptr,
                        const int *plan_br = nullptr, bool copied = false)
{
    const Proposal pr = draw_proposal(p.seed, r + p.rep_rng, step, p.n_mol);
    if (p.n_flex && p.mol_flex[pr.mol]) {
        const int kind = p.n_mol > 1 ? mulhi_n(pr.kind_word, 4) : 1 + mulhi_n(pr.kind_word, 3);
        if (kind >= 2) {
            propose_flexible(p, r, step, pr, kind, plan, plan_br, copied);
            return;
        }
        Proposal rigid = pr;
        rigid.rotate = kind;
        const Move mv = make_move(rigid, p.center[(size_t)r * p.n_mol + pr.mol], p.tstep[r], p.rcos[r], p.rsin[r],
                                  p.box);
        const int i_beg = p.mol_off[pr.mol], i_end = p.mol_off[pr.mol + 1];
        const float4 *rel = p.rel + (size_t)r * p.stride;
        for (int i = i_beg + threadIdx.x; i < i_end; i += kWG) {
            float4 rn, pn;
            move_residue(mv, rel[i], rn, pn);
            p.prop_pos[(size_t)r * p.max_n + (i - i_beg)] = pn;
            p.prop_rel[(size_t)r * p.max_n + (i - i_beg)] = rn;
        }
        __syncthreads();  // the proposal is in place
        const double bonded = proposal_bonded(p, r, pr.mol, -1);
        if (threadIdx.x == 0) {
            p.prop_mol[r] = pr.mol;
            p.prop_nib[r] = (i_end - i_beg + kIBlock - 1) / kIBlock;
            p.prop_n[r] = i_end - i_beg;
            p.prop_center[r] = make_float4(mv.c[0], mv.c[1], mv.c[2], 0.0f);
            // a rigid-body move keeps the radius; its rounded coordinates are
            // evaluated in full
            record_proposal(p, r, pr.mol, i_beg, i_end - i_beg, kMovedAll, 0, p.radius[(size_t)r * p.n_mol + pr.mol],
                            bonded);
        }
        return;
    }
    const Move mv = make_move(pr, p.center[(size_t)r * p.n_mol + pr.mol], p.tstep[r], p.rcos[r], p.rsin[r], p.box);
    const int i_beg = p.mol_off[pr.mol], i_end = p.mol_off[pr.mol + 1];
    const float4 *rel = p.rel + (size_t)r * p.stride;
    for (int i = i_beg + threadIdx.x; i < i_end; i += kWG) {
        float4 rn, pn;
        move_residue(mv, rel[i], rn, pn);
        p.prop_pos[(size_t)r * p.max_n + (i - i_beg)] = pn;
        p.prop_rel[(size_t)r * p.max_n + (i - i_beg)] = rn;
    }
    if (threadIdx.x == 0) {
        p.prop_mol[r] = pr.mol;
        p.prop_nib[r] = (i_end - i_beg + kIBlock - 1) / kIBlock;
        p.prop_n[r] = i_end - i_beg;
        p.prop_center[r] = make_float4(mv.c[0], mv.c[1], mv.c[2], 0.0f);
        record_proposal(p, r, pr.mol, i_beg, i_end - i_beg, kMovedNone, 0, p.radius[(size_t)r * p.n_mol + pr.mol],
                        0.0);
    }
}

__device__ void stage_one(const PairParams &p, int r);

// MC step, phase 2 (one workgroup per replica): Metropolis on the finished
// delta-E row (src/Replica.cpp:385-410, delta and exponent in FP64), apply
// the proposal (src/Molecule.cpp:468-504), update the matrix row and column
// and the running total, clear the row, then lay out the proposal of the
// next step.  With propose_only the first proposal is laid out.
__global__ __launch_bounds__(kWG) void mc_finish_kernel(PairParams p, int propose_only)
{
    __shared__ double red[kWavesPerWG];
    __shared__ int accept_flag;
    __shared__ FlexPlan sh_plan;
    __shared__ int sh_br[kWG];
    const int r = blockIdx.x;
    const int M = p.n_mol;
    const int64_t step = p.steps[r];
    // the next proposal's molecule and move kind follow from its random
    // words alone (src/Molecule.cpp:451-466, src/Replica.cpp:309-334)
    const int64_t next = propose_only ? step : step + 1;
    const Proposal nx = draw_proposal(p.seed, r + p.rep_rng, next, M);
    const bool next_flex = p.n_flex && p.mol_flex[nx.mol];
    const int kind2 = next_flex ? (M > 1 ? mulhi_n(nx.kind_word, 4) : 1 + mulhi_n(nx.kind_word, 3)) : 0;
    bool planned = false, copied = false;
    if (!propose_only) {
        // every load of the decision and of applying it is issued first
        // (the proposal header names the molecule and its residues); they
        // complete while waves 1-3 resolve the next move's topology
        const int t = threadIdx.x;
        const int4 hd = p.prop_hdr[r];
        const int m = hd.x, i_beg = hd.y, n_m = hd.z;
        const int lane = t & (kWave - 1), wave = t / kWave;
        double *row = p.rowbuf + (size_t)r * M;
        unsigned long long *fix = p.rowfix + (size_t)r * M;
        double *mat = p.emat + (size_t)r * M * M;
        float4 *pos = p.pos + (size_t)r * p.stride + i_beg;
        float4 *rel = p.rel + (size_t)r * p.stride + i_beg;
        float4 *pp = p.prop_pos + (size_t)r * p.max_n;
        float4 *pq = p.prop_rel + (size_t)r * p.max_n;
        double rv = 0.0, mv = 0.0;
        unsigned long long fv = 0ull;
        if (t < M) {
            rv = row[t];
            fv = fix[t];
            mv = mat[(size_t)m * M + t];
        }
        float4 a_pp = make_float4(0.f, 0.f, 0.f, 0.f), a_pq = a_pp, a_pos = a_pp, a_rel = a_pp;
        if (t < n_m) {
            a_pp = pp[t];
            a_pq = pq[t];
            a_pos = pos[t];
            a_rel = rel[t];
        }
        long long *cnt = p.counters + 3 * (size_t)r;
        double temp = 0.0, energy = 0.0, dS = 0.0;
        long long c0 = 0, c1 = 0, c2 = 0;
        int s1 = 1, flex_m = 0;
        if (t == 0) {
            temp = p.temperature[r];
            energy = p.energy[r];
            c0 = cnt[0];
            c1 = cnt[1];
            c2 = cnt[2];
            flex_m = p.internal && p.n_flex && p.mol_flex[m];
            if (p.da) {
                s1 = p.s1[r];
                dS = p.s_dS[r];
            }
        }
        // waves 1-3 resolve the topology of the next step's flexible move
        // (fixed whatever the decision)
        planned = kind2 >= 2;
        if (planned && t >= kWave) {
            FlexPlan fp;
            plan_flexible(p, r, next, nx.mol, kind2, fp);
            if (t == kWave) sh_plan = fp;
            if (kind2 == 2)
                for (int k = t - kWave; k < min(fp.nb, kWG); k += kWG - kWave) sh_br[k] = p.br[fp.b0 + k];
        }
        // dE = sum over J of row_J - matrix_mJ; the diagonal entry is the
        // moved molecule's internal energy when it is flexible (0 in both
        // the row and the matrix otherwise).  Row entries are completed with
        // their fixed-point part (the same thread reads entry J again below).
        double d = 0.0;
        if (t < M) {
            if (fv) {
                rv += (double)(long long)fv * kFixInv;
                row[t] = rv;
                fix[t] = 0ull;
            }
            d = rv - mv;
        }
        for (int J = t + kWG; J < M; J += kWG) {
            double v = row[J];
            const unsigned long long f = fix[J];
            if (f) {
                v += (double)(long long)f * kFixInv;
                row[J] = v;
                fix[J] = 0ull;
            }
            d += v - mat[(size_t)m * M + J];
        }
        // with at most 64 molecules every entry is wave 0's: thread 0 takes
        // the decision from its wave's sum without a barrier, so waves 1-3
        // meet wave 0 only once, after the decision (the other waves' sums
        // are zero, so the value is the same)
        const bool one_wave = M <= kWave;
        if (!one_wave || wave == 0) d = wave_total(d);
        if (!one_wave) {
            if (lane == 0) red[wave] = d;
            __syncthreads();
        }
        if (t == 0) {
            double delta = 0.0;
            if (one_wave)
                delta += d;
            else
                for (int w = 0; w < kWavesPerWG; ++w) delta += red[w];
            const Proposal pr = draw_proposal(p.seed, r + p.rep_rng, step, M);
            int acc = 0;
            if (p.da && !s1) {
                // rejected at stage one: no exact row was evaluated
                p.s1_rej[r] += 1;
            } else {
                // plain Metropolis (src/Replica.cpp:385-410), or delayed
                // acceptance stage two with min(1, exp(-(dE - dS) / kT))
                // (Christen & Fox 2005)
                const double d2 = p.da ? delta - dS : delta;
                if (d2 < 0.0) {
                    acc = 1;
                    cnt[0] = c0 + 1;
                } else if (pr.u < exp(-d2 * 4184.0 / ((double)kRgas * temp))) {
                    acc = 1;
                    cnt[1] = c1 + 1;
                } else {
                    cnt[2] = c2 + 1;
                }
                const unsigned long long nn = (unsigned long long)n_m;
                unsigned long long pairs = nn * ((unsigned long long)p.n_res - nn);
                if (flex_m) pairs += nn * (nn - 1) / 2;
                atomicAdd(p.pair_evals, pairs);
            }
            if (acc) p.energy[r] = energy + delta;
            accept_flag = acc;
        }
        __syncthreads();
        if (accept_flag) {
            // apply (src/Molecule.cpp:468-504): matrix row and column, the
            // molecule's coordinates, centre and radius
            if (t < M) {
                mat[(size_t)m * M + t] = rv;
                mat[(size_t)t * M + m] = rv;
            }
            for (int J = t + kWG; J < M; J += kWG) {
                const double v = row[J];
                mat[(size_t)m * M + J] = v;
                mat[(size_t)J * M + m] = v;
            }
            if (t < n_m) {
                pos[t] = a_pp;
                rel[t] = a_pq;
            }
            for (int k = t + kWG; k < n_m; k += kWG) {
                pos[k] = pp[k];
                rel[k] = pq[k];
            }
            if (t == 0) {
                p.center[(size_t)r * M + m] = p.prop_center[r];
                p.radius[(size_t)r * M + m] = p.prop_radius[r];
            }
            if (p.da) {
                float *sm = p.smat + (size_t)r * M * M;
                for (int J = t; J < M; J += kWG) {
                    if (J == m) continue;
                    const float v = p.srow[(size_t)r * M + J];
                    sm[(size_t)m * M + J] = v;
                    sm[(size_t)J * M + m] = v;
                }
                if (t < 9) p.frames[((size_t)r * M + m) * 9 + t] = p.prop_frame[(size_t)r * 9 + t];
            }
        } else if (planned && nx.mol == m) {
            // rejected, and the next flexible move is of the same molecule:
            // its proposal starts from the current coordinates
            if (t < n_m) {
                pp[t] = a_pos;
                pq[t] = a_rel;
            }
            for (int k = t + kWG; k < n_m; k += kWG) {
                pp[k] = pos[k];
                pq[k] = rel[k];
            }
        }
        // (accepted, the proposal buffers already hold the new coordinates)
        copied = planned && nx.mol == m;
        // (each row entry is read, applied and cleared by one thread)
        for (int J = t; J < M; J += kWG) row[J] = 0.0;
        __syncthreads();  // the new centre and relative positions are in place
    }
    if (!propose_only && threadIdx.x == 0) p.steps[r] = next;
    propose(p, r, next, planned ? &sh_plan : nullptr, planned ? sh_br : nullptr, copied);
    if (p.da && !propose_only) {
        __syncthreads();  // the proposal is laid out
        stage_one(p, r);
    }
}

// MC step, phase 2 for systems without flexible molecules and without
// delayed acceptance (one workgroup per replica): the same Metropolis,
// apply and next proposal as mc_finish_kernel, with every global load the
// step needs issued at the start.  The next step's molecule is known from
// its Philox counter before the decision, so its relative positions and
// centre are read together with the row, the matrix row and the pending
// proposal; they are read again only when the next molecule is the one
// just moved and the move was accepted.
__global__ __launch_bounds__(kWG) void mc_finish_rigid_kernel(PairParams p)
{
#pragma clang fp contract(off)
    __shared__ double red[kWavesPerWG];
    const int r = blockIdx.x, t = threadIdx.x;
    const int M = p.n_mol;
    const int64_t step = p.steps[r];
    const int m = p.prop_mol[r];
    const int i_beg = p.mol_off[m], n_m = p.mol_off[m + 1] - i_beg;
    const Proposal nx = draw_proposal(p.seed, r + p.rep_rng, step + 1, M);
    const int m2 = nx.mol;
    const int j_beg = p.mol_off[m2], n_2 = p.mol_off[m2 + 1] - j_beg;
    double *row = p.rowbuf + (size_t)r * M;
    unsigned long long *fix = p.rowfix + (size_t)r * M;
    double *mat = p.emat + (size_t)r * M * M;
    float4 *pos = p.pos + (size_t)r * p.stride;
    float4 *rel = p.rel + (size_t)r * p.stride;
    const float4 *pp = p.prop_pos + (size_t)r * p.max_n;
    const float4 *pq = p.prop_rel + (size_t)r * p.max_n;
    // loads of the whole step (the first element of each per-thread loop)
    double rv = 0.0, mv_old = 0.0;
    unsigned long long fv = 0ull;
    if (t < M) {
        rv = row[t];
        fv = fix[t];
        mv_old = mat[(size_t)m * M + t];
    }
    float4 a_pos = make_float4(0.f, 0.f, 0.f, 0.f), a_rel = a_pos, n_rel = a_pos;
    if (t < n_m) {
        a_pos = pp[t];
        a_rel = pq[t];
    }
    if (t < n_2) n_rel = rel[j_beg + t];
    float4 c2 = p.center[(size_t)r * M + m2];
    const float4 c_new = p.prop_center[r];
    const float tstep = p.tstep[r];
    const double rcos = p.rcos[r], rsin = p.rsin[r];
    const double temp = p.temperature[r];
    // Metropolis (src/Replica.cpp:385-410, delta and exponent in FP64)
    double d = 0.0;
    if (t < M) {
        if (fv) rv += (double)(long long)fv * kFixInv;
        d = rv - mv_old;
    }
    for (int J = t + kWG; J < M; J += kWG) {
        double v = row[J];
        const unsigned long long f = fix[J];
        if (f) v += (double)(long long)f * kFixInv;
        d += v - mat[(size_t)m * M + J];
    }
    d = wave_total(d);
    if ((t & (kWave - 1)) == 0) red[t / kWave] = d;
    // one barrier: the wave sums are in LDS and every wave has read the
    // step's inputs (nothing read below is written by another thread)
    __syncthreads();
    // every thread takes the same decision (same inputs, same operations);
    // thread 0 records it
    double delta = 0.0;
    for (int w = 0; w < kWavesPerWG; ++w) delta += red[w];
    const Proposal pr = draw_proposal(p.seed, r + p.rep_rng, step, M);
    const int kind = delta < 0.0 ? 0 : pr.u < exp(-delta * 4184.0 / ((double)kRgas * temp)) ? 1 : 2;
    const int acc = kind < 2;
    if (t == 0) {
        p.counters[3 * (size_t)r + kind] += 1;
        const unsigned long long nm = (unsigned long long)n_m;
        atomicAdd(p.pair_evals, nm * ((unsigned long long)p.n_res - nm));
        if (acc) p.energy[r] += delta;
        p.steps[r] = step + 1;
    }
    if (acc) {
        // apply (src/Molecule.cpp:468-504) and the matrix row and column
        if (t < M) {
            mat[(size_t)m * M + t] = rv;
            mat[(size_t)t * M + m] = rv;
        }
        for (int J = t + kWG; J < M; J += kWG) {
            double v = row[J];
            const unsigned long long f = fix[J];
            if (f) v += (double)(long long)f * kFixInv;
            mat[(size_t)m * M + J] = v;
            mat[(size_t)J * M + m] = v;
        }
        if (t < n_m) {
            pos[i_beg + t] = a_pos;
            rel[i_beg + t] = a_rel;
        }
        for (int k = t + kWG; k < n_m; k += kWG) {
            pos[i_beg + k] = pp[k];
            rel[i_beg + k] = pq[k];
        }
        if (t == 0) p.center[(size_t)r * M + m] = c_new;
    }
    // (each row entry, and each proposal record, is read, applied and
    // replaced by one thread: no barrier before they are cleared)
    for (int J = t; J < M; J += kWG) {
        row[J] = 0.0;
        fix[J] = 0ull;
    }
    // the next proposal (src/Molecule.cpp:146-266); when it moves the
    // molecule just accepted, it starts from the applied coordinates, which
    // this thread holds (residues past the first kWG are re-read below, after
    // this thread wrote them)
    if (acc && m2 == m) {
        n_rel = a_rel;
        c2 = c_new;
    }
    const Move mv = make_move(nx, c2, tstep, rcos, rsin, p.box);
    float4 *qp = p.prop_pos + (size_t)r * p.max_n;
    float4 *qq = p.prop_rel + (size_t)r * p.max_n;
    if (t < n_2) {
        float4 rn, pn;
        move_residue(mv, n_rel, rn, pn);
        qp[t] = pn;
        qq[t] = rn;
    }
    for (int k = t + kWG; k < n_2; k += kWG) {
        float4 rn, pn;
        move_residue(mv, rel[j_beg + k], rn, pn);
        qp[k] = pn;
        qq[k] = rn;
    }
    if (t == 0) {
        p.prop_mol[r] = m2;
        p.prop_nib[r] = (n_2 + kIBlock - 1) / kIBlock;
        p.prop_n[r] = n_2;
        p.prop_center[r] = make_float4(mv.c[0], mv.c[1], mv.c[2], 0.0f);
        record_proposal(p, r, m2, j_beg, n_2, kMovedNone, 0, p.radius[(size_t)r * M + m2], 0.0);
    }
}

// Value of window row entry J of slot s of replica r: the fixed-point sum
// plus the FP64 part (partials too large for the fixed point).
__device__ __forceinline__ double window_row(const PairParams &p, int r, int s, int J)
{
    const size_t at = ((size_t)(r << p.win_log2) + s) * (p.n_mol + (1 << p.win_log2)) + J;
    double v = p.wrowbuf[at];
    const unsigned long long f = p.wrowfix[at];
    if (f) v += (double)(long long)f * kFixInv;
    return v;
}

__device__ __forceinline__ WinDraw win_draw(const PairParams &p, int r, int rr, int64_t step)
{
    const Proposal pr = draw_proposal(p.seed, rr, step, p.n_mol);
    WinDraw d;
    d.mol = pr.mol;
    d.beg = p.mol_off[pr.mol];
    d.n = p.mol_off[pr.mol + 1] - d.beg;
    move_direction(pr, p.tstep[r], p.rcos[r], p.rsin[r], d);
    d.u = pr.u;
    return d;
}

// One workgroup per replica.  begin_steps >= 0: the start of a run of
// begin_steps MC steps: sets the run's end and lays out the first window
// (slot 0 is the pending proposal; slots 1.. are laid out from the current
// state).  begin_steps < 0: takes the decisions of the current window in
// step order (src/Replica.cpp:385-410, FP64 delta and exponent), applies
// the accepted moves and the matrix rows and columns, and lays out the
// next window (slot 0 always, so that a proposal is pending when the run
// ends).  A replica that has reached the run's end does nothing.
//
// Decision of slot s: with D_s = sum_J (row_s[J] - mat[m_s][J]) over the
// window-start state, and for each earlier slot q that was accepted the
// replacement of its entry row_s[m_q] - mat[m_s][m_q] by
// X_sq - row_q[m_s] (X_sq the energy between the two proposals, row_q[m_s]
// the energy between q's proposal and m_s as stored):
//   delta_s = D_s + sum_{q < s accepted} (X_sq - row_q[m_s] - row_s[m_q] + mat[m_s][m_q]).
// Metropolis accepts when u < exp(-delta / RT); the test is taken on
// log(u) (computed while the loads are in flight) unless the two sides are
// within 1e-12 of each other, where the exponential decides as written.
//
// The kernel is a chain of latencies, so every load is issued before the
// decisions and every store after them, in two levels: (1) the window's
// headers, uniforms and rows and the next window's draws (laid out by the
// previous launch), (2) what depends on their molecules: the matrix
// entries of the delta-E sums, the correction terms, this window's
// proposed relative positions (for the apply) and the next molecules'
// current relative positions and centres (for the next proposals), staged
// in LDS.  A proposed position is the proposed centre plus the proposed
// relative position (the order of move_residue's float additions), so only
// relative positions are read.  The code is kept compact (one site per
// loop): every launch runs it from a cold instruction cache.
constexpr int kWinStage = 1536;  // residues staged in LDS per kind (this window's, the next window's)
// The decision kernel's workgroup: 16 waves, so that the residues of a
// window are spread thin and every slot has a wave of its own.
constexpr int kWinWaves = 16;
constexpr int kWinWG = kWinWaves * kWave;
constexpr int kWinIt = (kMaxWin + kWinWaves - 1) / kWinWaves;  // slots per wave
__global__ __launch_bounds__(kWinWG) void mc_window_kernel(PairParams p, int begin_steps)
{
#pragma clang fp contract(off)
    __shared__ double s_D[kMaxWin];
    __shared__ double s_corr[kMaxWin * kMaxWin];
    __shared__ double s_X[kMaxWin * kMaxWin];
    __shared__ double s_u[kMaxWin], s_lu[kMaxWin];
    __shared__ int4 s_hd[kMaxWin];
    __shared__ float4 s_cn[kMaxWin];
    __shared__ WinDraw s_nd[kMaxWin];      // the next window's draws
    __shared__ float s_nrad[kMaxWin];
    __shared__ Move s_mv[2][kMaxWin];      // from the stored centre, and from the proposed centre of
                                           // the slot of this window that moves the same molecule
    __shared__ RotTerms s_rt[kMaxWin];     // the next slots' rotation terms
    __shared__ unsigned s_accm;
    __shared__ float4 s_ar[kWinStage];     // this window's proposed relative positions
    __shared__ float4 s_nr[kWinStage];     // the next molecules' stored relative positions
    const int r = blockIdx.x, t = threadIdx.x;
    const int lane = t & (kWave - 1), wave = t / kWave;
    const int lw = p.win_log2, W = 1 << lw;
    const int M = p.n_mol, N = p.n_res;
    const bool begin = begin_steps >= 0;
    // (every load of the first level is issued before a replica that has
    // reached its run's end leaves; none of them depends on the window's
    // length)
    const int64_t step = p.steps[r];
    const int act = begin ? 0 : p.wlen[r];
    const int64_t stop = begin ? step + begin_steps : p.wstop[r];
    const int64_t next = step + act;
    const int rr = r + p.rep_rng;
    const int first = begin ? 1 : 0;
    float4 *pos = p.pos + (size_t)r * p.stride;
    float4 *rel = p.rel + (size_t)r * p.stride;
    double *mat = p.emat + (size_t)r * M * M;
    const double temp = p.temperature[r];
    const double beta = 4184.0 / ((double)kRgas * temp);
    // ---- level 1: the window and the next window's draws
    double e_old = 0.0;
    long long c0 = 0, c1 = 0, c2 = 0;
    if (t == 0) {
        e_old = p.energy[r];
        c0 = p.counters[3 * (size_t)r];
        c1 = p.counters[3 * (size_t)r + 1];
        c2 = p.counters[3 * (size_t)r + 2];
    }
    if (t < W) {
        s_hd[t] = p.whdr[(r << lw) + t];
        s_cn[t] = p.wcen[(r << lw) + t];
        s_u[t] = p.wu[(r << lw) + t];
    } else if (t >= kWave && t < kWave + W) {
        const int k = t - kWave;
        s_nd[k] = begin ? win_draw(p, r, rr, step + k) : p.wd[(r << lw) + k];
    }
    // rows: wave w takes slots w and w + 4, its lanes J, J + 64, ... (the
    // values stay in registers for the matrix update)
    constexpr int kRowRegs = 4;
    double rv[kWinIt][kRowRegs];
#pragma unroll
    for (int it = 0; it < kWinIt; ++it) {
        const int s = wave + it * kWinWaves;
#pragma unroll
        for (int u = 0; u < kRowRegs; ++u) {
            const int J = lane + u * kWave;
            rv[it][u] = s < W && J < M ? window_row(p, r, s, J) : 0.0;
        }
    }
    if (!begin && act == 0) return;
    __syncthreads();
    // ---- level 2
    const int len = window_len([&](int k) { return s_nd[k].mol; }, next, stop, W);
    const int n_lay = begin ? len : max(len, 1);
    // residue prefixes: this window's slots, the next window's laid-out slots
    int cum[kMaxWin + 1], ncum[kMaxWin + 1];
    cum[0] = ncum[0] = 0;
#pragma unroll
    for (int s = 0; s < kMaxWin; ++s) {
        cum[s + 1] = cum[s] + (s < act ? s_hd[s].z : 0);
        ncum[s + 1] = ncum[s] + (s >= first && s < n_lay ? s_nd[s].n : 0);
    }
    const int n_items = cum[kMaxWin], n_next = ncum[kMaxWin];
    // the slot of this window that moves next slot j's molecule, or -1
    auto slot_moving = [&](int j) {
        int a = -1;
        for (int q = 0; q < act; ++q)
            if (s_hd[q].x == s_nd[j].mol) a = q;
        return a;
    };
    // slot of item i in a prefix, and a prefix entry (selects, so that the
    // prefixes stay in registers)
    auto slot_of = [&](const int (&c)[kMaxWin + 1], int i) {
        int s = 0;
#pragma unroll
        for (int q = 1; q < kMaxWin; ++q) s += i >= c[q] ? 1 : 0;
        return s;
    };
    auto at = [&](const int (&c)[kMaxWin + 1], int s) {
        int v = 0;
#pragma unroll
        for (int q = 0; q <= kMaxWin; ++q) v = q == s ? c[q] : v;
        return v;
    };
    if (t >= kWave + kMaxWin && t < kWave + kMaxWin + act) {
        const int k = t - kWave - kMaxWin;
        s_lu[k] = log(s_u[k]);
    } else if (begin && t == kWave + 2 * kMaxWin) {
        // slot 0 of a run's first window: the pending proposal
        const int4 h = p.prop_hdr[r];
        const float4 c = p.prop_center[r];
        p.whdr[r << lw] = h;
        p.wcen[r << lw] = make_float4(c.x, c.y, c.z, __int_as_float(0));
    }
    if (t >= 2 * kWave && t < 2 * kWave + W * W) {
        const int s = (t - 2 * kWave) >> lw, q = (t - 2 * kWave) & (W - 1);
        if (q < s && s < act) {
            const int ms = s_hd[s].x, mq = s_hd[q].x;
            const double X = window_row(p, r, s, M + q);
            const double a = window_row(p, r, q, ms);
            const double b = window_row(p, r, s, mq);
            const double c = mat[(size_t)ms * M + mq];
            s_X[(s << lw) + q] = X;
            s_corr[(s << lw) + q] = X - a - (b - c);
        }
    }
    // D_s
#pragma unroll
    for (int it = 0; it < kWinIt; ++it) {
        const int s = wave + it * kWinWaves;
        if (s >= act) continue;
        const double *mrow = mat + (size_t)s_hd[s].x * M;
        double d = 0.0;
#pragma unroll
        for (int u = 0; u < kRowRegs; ++u) {
            const int J = lane + u * kWave;
            if (J < M) d += rv[it][u] - mrow[J];
        }
        for (int J = lane + kRowRegs * kWave; J < M; J += kWave) d += window_row(p, r, s, J) - mrow[J];
        d = wave_total(d);
        if (lane == 0) s_D[s] = d;
    }
    // staged: this window's proposed relative positions, the next
    // molecules' stored ones (three per thread per pass, loads first)
    for (int base = 0; base < kWinStage && (base < n_items || base < n_next); base += 3 * kWinWG) {
        auto load_a = [&](int i) {
            const int s = slot_of(cum, i);
            return i < n_items && i < kWinStage ? slot_rel(p, r, s)[i - at(cum, s)] : make_float4(0.f, 0.f, 0.f, 0.f);
        };
        auto load_b = [&](int i) {
            const int j = slot_of(ncum, i);
            return i < n_next && i < kWinStage ? rel[s_nd[j].beg + i - at(ncum, j)] : make_float4(0.f, 0.f, 0.f, 0.f);
        };
        const int i0 = base + t, i1 = i0 + kWinWG, i2 = i1 + kWinWG;
        const float4 a0 = load_a(i0), a1 = load_a(i1), a2 = load_a(i2);
        const float4 b0 = load_b(i0), b1 = load_b(i1), b2 = load_b(i2);
        if (i0 < kWinStage) {
            s_ar[i0] = a0;
            s_nr[i0] = b0;
        }
        if (i1 < kWinStage) {
            s_ar[i1] = a1;
            s_nr[i1] = b1;
        }
        if (i2 < kWinStage) {
            s_ar[i2] = a2;
            s_nr[i2] = b2;
        }
    }
    __syncthreads();
    // ---- decisions in step order (wave 0: lane l holds slot l's delta-E,
    // its uniform and its correction terms; each acceptance adds its
    // column of corrections to the later slots), while wave 2 forms the
    // next slots' moves and wave 3 draws the steps after the next window;
    // then every store
    if (wave == 2) {
        if (lane < 2 * kMaxWin) {
            // the next slots' moves: from the stored centre (var 0), and,
            // when this window moves the same molecule, from its proposed
            // centre
            const int j = lane & (kMaxWin - 1), var = lane / kMaxWin;
            const int a = j < n_lay ? slot_moving(j) : -1;
            if (j >= first && j < n_lay && (var == 0 || a >= 0)) {
                const WinDraw &d = s_nd[j];
                const float4 c2 = var == 0 ? p.center[(size_t)r * M + d.mol] : s_cn[a];
                if (var == 0) s_nrad[j] = p.radius[(size_t)r * M + d.mol];
                const Move mv = move_at(d, c2, p.box);
                s_mv[var][j] = mv;
                if (var == 0) s_rt[j] = rot_terms(mv);
            }
        }
    } else if (wave == 3) {
        if (lane < kMaxWin) {
            // the draws after the next window (for the launch after this
            // one), and the next window's uniforms
            const int k = lane;
            if (k < W) p.wd[(r << lw) + k] = win_draw(p, r, rr, next + len + k);
            if (k < len) p.wu[(r << lw) + k] = s_nd[k].u;
        }
    } else if (wave == 0) {
        double dl = 0.0, lul = 0.0, ul = 1.0;
        double cl[kMaxWin];
        if (lane < act) {
            dl = s_D[lane];
            lul = s_lu[lane];
            ul = s_u[lane];
        }
#pragma unroll
        for (int q = 0; q < kMaxWin; ++q) cl[q] = lane < act && q < lane ? s_corr[(lane << lw) + q] : 0.0;
        unsigned accm = 0u;
        double e = e_old;
        unsigned long long pairs = 0ull;
        for (int s = 0; s < act; ++s) {
            const double d = __longlong_as_double(readlane_b64(__double_as_longlong(dl), s));
            const double lu = __longlong_as_double(readlane_b64(__double_as_longlong(lul), s));
            const double u = __longlong_as_double(readlane_b64(__double_as_longlong(ul), s));
            int kind;
            if (d < 0.0) {
                kind = 0;
            } else {
                const double x = -d * beta;
                const double gap = x - lu;
                if (u != 0.0 && fabs(gap) > 1e-12 * (1.0 + fabs(x)))
                    kind = gap > 0.0 ? 1 : 2;
                else
                    kind = u < exp(-d * 4184.0 / ((double)kRgas * temp)) ? 1 : 2;
            }
            if (kind < 2) {
                accm |= 1u << s;
                e += d;
                double c = 0.0;
#pragma unroll
                for (int q = 0; q < kMaxWin; ++q) c = q == s ? cl[q] : c;
                dl += c;
            }
            c0 += kind == 0;
            c1 += kind == 1;
            c2 += kind == 2;
            const unsigned long long nm = (unsigned long long)s_hd[s].z;
            pairs += nm * ((unsigned long long)N - nm);
        }
        if (lane == 0) s_accm = accm;
        if (lane == 0 && act > 0) {
            p.energy[r] = e;
            p.counters[3 * (size_t)r] = c0;
            p.counters[3 * (size_t)r + 1] = c1;
            p.counters[3 * (size_t)r + 2] = c2;
            p.steps[r] = next;
            atomicAdd(p.pair_evals, pairs);
        }
        if (lane == 0) {
            p.wlen[r] = len;
            if (begin) p.wstop[r] = stop;
        }
    }
    __syncthreads();
    const unsigned accm = s_accm;
    // apply (src/Molecule.cpp:468-504): the accepted slots' coordinates
    // (residues past the staged ones are read now, before the barrier that
    // precedes the next window's writes into the slot buffers)
    for (int i = t; i < n_items; i += kWinWG) {
        const int s = slot_of(cum, i);
        if (accm >> s & 1u) {
            const int k = i - at(cum, s);
            const float4 q = i < kWinStage ? s_ar[i] : slot_rel(p, r, s)[k], c = s_cn[s];
            rel[s_hd[s].y + k] = q;
            pos[s_hd[s].y + k] = make_float4(c.x + q.x, c.y + q.y, c.z + q.z, q.w);
        }
    }
    if (t < act && (accm >> t & 1u))
        p.center[(size_t)r * M + s_hd[t].x] = make_float4(s_cn[t].x, s_cn[t].y, s_cn[t].z, 0.0f);
    // the matrix rows and columns of the accepted slots; an entry between
    // two accepted slots is the energy between their proposals
#pragma unroll
    for (int it = 0; it < kWinIt; ++it) {
        const int s = wave + it * kWinWaves;
        if (s >= act || !(accm >> s & 1u)) continue;
        const int ms = s_hd[s].x;
        for (int J = lane; J < M; J += kWave) {
            const int u = J / kWave;
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < kRowRegs; ++q)
                if (u == q) v = rv[it][q];
            if (u >= kRowRegs) v = window_row(p, r, s, J);
            for (int q = 0; q < act; ++q)
                if (q != s && (accm >> q & 1u) && s_hd[q].x == J) v = q < s ? s_X[(s << lw) + q] : s_X[(q << lw) + s];
            mat[(size_t)ms * M + J] = v;
            mat[(size_t)J * M + ms] = v;
        }
    }
    // the rows are cleared for the next window
    for (int k = t; k < act * (M + W); k += kWinWG) {
        const size_t at = (size_t)(r << lw) * (M + W) + k;
        p.wrowfix[at] = 0ull;
        p.wrowbuf[at] = 0.0;
    }
    // ---- the next window (src/Molecule.cpp:146-266), from the state after
    // the apply: a next molecule that this window moved and accepted starts
    // from that proposal.  Past the staged residues the relative positions
    // after the apply are read, behind a barrier.
    if (n_items > kWinStage || n_next > kWinStage) __syncthreads();
    for (int i = t; i < n_next; i += kWinWG) {
        const int j = slot_of(ncum, i), k = i - at(ncum, j);
        const int a = slot_moving(j);
        const bool from_slot = a >= 0 && (accm >> a & 1u);
        const int ia = from_slot ? at(cum, a) + k : 0;
        float4 q;
        if (from_slot && ia < kWinStage)
            q = s_ar[ia];
        else if (!from_slot && i < kWinStage)
            q = s_nr[i];
        else
            q = rel[s_nd[j].beg + k];
        float4 rn, pn;
        move_residue_terms(s_mv[from_slot ? 1 : 0][j], s_rt[j], q, rn, pn);
        slot_pos(p, r, j)[k] = pn;
        slot_rel(p, r, j)[k] = rn;
    }
    if (t < n_lay && t >= first) {
        const int j = t;
        const int a = slot_moving(j);
        const Move &mv = s_mv[a >= 0 && (accm >> a & 1u) ? 1 : 0][j];
        const int m2 = s_nd[j].mol, j_beg = s_nd[j].beg, n_2 = s_nd[j].n;
        const float rad = s_nrad[j];
        p.whdr[(r << lw) + j] = make_int4(m2, j_beg, n_2, __float_as_int(rad));
        p.wcen[(r << lw) + j] = make_float4(mv.c[0], mv.c[1], mv.c[2], __int_as_float(at(ncum, j) + (first ? s_nd[0].n : 0)));
        if (j == 0) {
            p.prop_mol[r] = m2;
            p.prop_nib[r] = (n_2 + kIBlock - 1) / kIBlock;
            p.prop_n[r] = n_2;
            p.prop_center[r] = make_float4(mv.c[0], mv.c[1], mv.c[2], 0.0f);
            record_proposal(p, r, m2, j_beg, n_2, kMovedNone, 0, rad, 0.0);
        }
    }
}

// ---------------------------------------------------------------------------
// Resident MC steps for small rigid systems (every residue of a replica in
// LDS): one workgroup per replica runs many MC steps in one launch.  The
// replica's positions, relative positions, centres and pending proposal
// live in LDS for the whole launch; a step evaluates the proposal's row
// against every other residue, decides, applies and lays out the next
// proposal with workgroup barriers only, so a step costs LDS latencies and
// a few thousand pair terms instead of two kernel launches.  Systems: no
// flexible molecules, no delayed acceptance, no repulsive crowding, at
// most kResidentRes residues and kResidentMol molecules.
//
// The row: a thread takes a j residue outside the moved molecule m against
// a range of m's proposed residues (the residues split into as many ranges
// as the workgroup's threads can take side by side), summing in Kahan FP32
// with the per-pair minimum image; each molecule J's lanes are summed in
// FP64 (a fixed DPP order) into the wave's slot for J, and the four waves'
// slots are added in wave order: the row does not depend on timing.  The decision,
// apply and next proposal are those of mc_finish_rigid_kernel.
constexpr int kResidentRes = 1024;
constexpr int kResidentMol = 256;

struct ResidentParams {
    int steps;  // MC steps of this launch
};

// Dynamic LDS of the resident kernel for n residues, m molecules and
// proposals of at most max_n residues.
__host__ __device__ inline size_t resident_lds_bytes(int n, int m, int max_n)
{
    return (size_t)16 * (2 * (size_t)n + 4 * (size_t)max_n + (size_t)m) + (size_t)8 * 2 * kWavesPerWG * m +
           (size_t)4 * ((size_t)n + (size_t)m + 1);
}

__global__ __launch_bounds__(kWG) void mc_resident_kernel(PairParams p, ResidentParams rp)
{
#pragma clang fp contract(off)
    __shared__ __attribute__((aligned(16))) char tab[kTableBytes];
    extern __shared__ __attribute__((aligned(16))) float4 res_lds[];
    const int r = blockIdx.x, t = threadIdx.x;
    const int lane = t & (kWave - 1), wave = t / kWave;
    const int N = p.n_res, M = p.n_mol;
    float4 *s_pos = res_lds;                  // N
    float4 *s_rel = s_pos + N;                // N
    float4 *s_pa = s_rel + N;                 // 2 x max_n: proposed positions (two slots)
    float4 *s_ra = s_pa + 2 * p.max_n;        // 2 x max_n: proposed relative positions
    float4 *s_cen = s_ra + 2 * p.max_n;       // M
    double *s_part = reinterpret_cast<double *>(s_cen + M);  // 2 x kWavesPerWG x M: the waves' row slots
    int *s_mol = reinterpret_cast<int *>(s_part + 2 * kWavesPerWG * M);  // N: molecule of each residue
    int *s_off = s_mol + N;                                              // M + 1: molecule offsets
    // the table (entry e = tj * 20 + ti, i = the proposal's residue, at row
    // ti, column tj, kcal/mol) and the replica's state
    {
        constexpr int kEntries = CGPPD_AA_COUNT * CGPPD_AA_COUNT;
        const int e0 = t, e1 = min(t + kWG, kEntries - 1);
        const Coef g0 = p.table[e0], g1 = p.table[e1];
        *reinterpret_cast<Coef *>(tab + (((e0 % CGPPD_AA_COUNT) << kRowShift) | ((e0 / CGPPD_AA_COUNT) << kColShift))) = g0;
        *reinterpret_cast<Coef *>(tab + (((e1 % CGPPD_AA_COUNT) << kRowShift) | ((e1 / CGPPD_AA_COUNT) << kColShift))) = g1;
    }
    for (int k = t; k < N; k += kWG) {
        s_pos[k] = p.pos[(size_t)r * p.stride + k];
        s_rel[k] = p.rel[(size_t)r * p.stride + k];
    }
    for (int k = t; k < M; k += kWG) s_cen[k] = p.center[(size_t)r * M + k];
    for (int k = t; k < N; k += kWG) s_mol[k] = p.mol_of[k];
    for (int k = t; k <= M; k += kWG) s_off[k] = p.mol_off[k];
    for (int k = t; k < kWavesPerWG * M; k += kWG) s_part[k] = 0.0;
    const int n_first = p.prop_n[r];
    for (int k = t; k < n_first; k += kWG) {
        s_pa[k] = p.prop_pos[(size_t)r * p.max_n + k];
        s_ra[k] = p.prop_rel[(size_t)r * p.max_n + k];
    }
    float4 pc = p.prop_center[r];  // the pending proposal's centre (every thread)
    int64_t step = p.steps[r];
    double energy = p.energy[r];
    long long cnt[3] = {0, 0, 0};
    unsigned long long pairs = 0;
    const float tstep = p.tstep[r];
    const double rcos = p.rcos[r], rsin = p.rsin[r];
    const double temp = p.temperature[r];
    double *mat = p.emat + (size_t)r * M * M;
    Proposal pr = draw_proposal(p.seed, r + p.rep_rng, step, M);
    __syncthreads();
    for (int s = 0; s < rp.steps; ++s) {
        // slot cur holds this step's proposal, the other one the next
        // step's; the row slots alternate the same way
        const int cur = s & 1;
        const float4 *prop = s_pa + cur * p.max_n, *prel = s_ra + cur * p.max_n;
        float4 *nprop = s_pa + (cur ^ 1) * p.max_n, *nprel = s_ra + (cur ^ 1) * p.max_n;
        double *part = s_part + cur * kWavesPerWG * M;
        const int m = pr.mol;
        const int i0 = s_off[m], n_m = s_off[m + 1] - i0;
        const int nj = N - n_m;
        // the matrix row of m, read now (before any wave can apply this
        // step's move to it) and used after the pair loop
        constexpr int kRowPerLane = kResidentMol / kWave;
        double mrow[kRowPerLane];
#pragma unroll
        for (int h = 0; h < kRowPerLane; ++h)
            mrow[h] = lane + h * kWave < M ? mat[(size_t)m * M + lane + h * kWave] : 0.0;
        // the next step's proposal, laid out now from the current state (it
        // holds unless this step moves the same molecule and is accepted)
        const Proposal nx = draw_proposal(p.seed, r + p.rep_rng, step + 1, M);
        const int m2 = nx.mol;
        const int j_beg = s_off[m2], n_2 = s_off[m2 + 1] - j_beg;
        const Move mv = make_move(nx, s_cen[m2], tstep, rcos, rsin, p.box);
        for (int k = t; k < n_2; k += kWG) {
            float4 rn, pn;
            move_residue(mv, s_rel[j_beg + k], rn, pn);
            nprop[k] = pn;
            nprel[k] = rn;
        }
        float4 npc = make_float4(mv.c[0], mv.c[1], mv.c[2], 0.0f);
        // work items q = (range si of the proposal's residues, j residue jp
        // outside m): S ranges, as many as the threads can take side by side
        const int S = nj >= kWG || nj == 0 ? 1 : kWG / nj;
        for (int q0 = wave * kWave; q0 < S * nj; q0 += kWG) {
            const int q = q0 + lane;
            const bool valid = q < S * nj;
            const int jp = valid ? q % nj : 0, si = valid ? q / nj : 0;
            const int ia = si * n_m / S, ib = (si + 1) * n_m / S;
            const int j = min(jp + (jp >= i0 ? n_m : 0), N - 1);
            const float4 rj = s_pos[j];
            const int J = s_mol[j];
            // row = the proposal residue's type (the same for every lane),
            // column = the j residue's: a wave's lanes read one 320-byte row
            // and do not collide in LDS banks
            const uint32_t jcol = lane_col(rj);
            float sum = 0.0f, comp = 0.0f;
            // four proposal residues at a time: their records and table
            // entries are read before any term is formed
            int i = valid ? ia : ib;
            for (; i + 4 <= ib; i += 4) {
                float4 pi[4];
                Coef cf[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) pi[u] = prop[i + u];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    cf[u] = *reinterpret_cast<const Coef *>(tab + ((__float_as_uint(pi[u].w) & kRowMask) | jcol));
                float g = 0.0f;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    float lj, dh;
                    pair_term<true, true>(rj, pi[u].x, pi[u].y, pi[u].z, cf[u], p.box, p.inv_box, lj, dh);
                    g += lj + dh;
                }
                kahan_add(sum, comp, g);
            }
            for (; i < ib; ++i) {
                const float4 pi = prop[i];
                const Coef cf = *reinterpret_cast<const Coef *>(tab + ((__float_as_uint(pi.w) & kRowMask) | jcol));
                float lj, dh;
                pair_term<true, true>(rj, pi.x, pi.y, pi.z, cf, p.box, p.inv_box, lj, dh);
                kahan_add(sum, comp, lj + dh);
            }
            const double v = valid ? (double)sum - (double)comp : 0.0;
            // the molecules of the wave's items, one at a time in lane order
            unsigned long long todo = __ballot(valid);
            while (todo) {
                const int Jc = __builtin_amdgcn_readlane(J, __builtin_ctzll(todo));
                const bool mine = valid && J == Jc;
                const double tt = wave_total(mine ? v : 0.0);
                if (lane == 0) part[wave * M + Jc] += tt;
                todo &= ~__ballot(mine);
            }
        }
        __syncthreads();
        // Metropolis (src/Replica.cpp:385-410, delta and exponent in FP64),
        // taken by every wave on the same values in the same order
        double d = 0.0;
#pragma unroll
        for (int h = 0; h < kRowPerLane; ++h) {
            const int J = lane + h * kWave;
            if (J < M) {
                double e = 0.0;
                for (int w = 0; w < kWavesPerWG; ++w) e += part[w * M + J];
                d += e - mrow[h];
            }
        }
        const double delta = wave_total(d);
        const int kind = delta < 0.0 ? 0 : pr.u < exp(-delta * 4184.0 / ((double)kRgas * temp)) ? 1 : 2;
        const bool acc = kind < 2;
        cnt[kind] += 1;
        pairs += (unsigned long long)n_m * (unsigned long long)nj;
        if (acc) {
            // apply (src/Molecule.cpp:468-504) and the matrix row and column
            energy += delta;
            for (int J = t; J < M; J += kWG) {
                double e = 0.0;
                for (int w = 0; w < kWavesPerWG; ++w) e += part[w * M + J];
                mat[(size_t)m * M + J] = e;
                mat[(size_t)J * M + m] = e;
            }
            for (int k = t; k < n_m; k += kWG) {
                s_pos[i0 + k] = prop[k];
                s_rel[i0 + k] = prel[k];
            }
            if (t == 0) s_cen[m] = pc;
            if (m2 == m) {
                // the next move starts from the accepted coordinates
                const Move mva = make_move(nx, pc, tstep, rcos, rsin, p.box);
                for (int k = t; k < n_2; k += kWG) {
                    float4 rn, pn;
                    move_residue(mva, prel[k], rn, pn);
                    nprop[k] = pn;
                    nprel[k] = rn;
                }
                npc = make_float4(mva.c[0], mva.c[1], mva.c[2], 0.0f);
            }
        }
        // the next step's row slots start at zero
        for (int k = t; k < kWavesPerWG * M; k += kWG) s_part[(cur ^ 1) * kWavesPerWG * M + k] = 0.0;
        __syncthreads();
        pr = nx;
        pc = npc;
        ++step;
    }
    // the replica's state back to memory, and the pending proposal in the
    // form the other MC kernels read
    const int last = rp.steps & 1;
    for (int k = t; k < N; k += kWG) {
        p.pos[(size_t)r * p.stride + k] = s_pos[k];
        p.rel[(size_t)r * p.stride + k] = s_rel[k];
    }
    for (int k = t; k < M; k += kWG) p.center[(size_t)r * M + k] = s_cen[k];
    const int m2 = pr.mol;
    const int j_beg = s_off[m2], n_2 = s_off[m2 + 1] - j_beg;
    for (int k = t; k < n_2; k += kWG) {
        p.prop_pos[(size_t)r * p.max_n + k] = s_pa[last * p.max_n + k];
        p.prop_rel[(size_t)r * p.max_n + k] = s_ra[last * p.max_n + k];
    }
    if (t == 0) {
        p.energy[r] = energy;
        long long *c = p.counters + 3 * (size_t)r;
        c[0] += cnt[0];
        c[1] += cnt[1];
        c[2] += cnt[2];
        p.steps[r] = step;
        atomicAdd(p.pair_evals, pairs);
        p.prop_mol[r] = m2;
        p.prop_nib[r] = (n_2 + kIBlock - 1) / kIBlock;
        p.prop_n[r] = n_2;
        p.prop_center[r] = pc;
        record_proposal(p, r, m2, j_beg, n_2, kMovedNone, 0, p.radius[(size_t)r * M + m2], 0.0);
    }
}

// After the scaled full-energy passes: mirror the upper triangle of each
// replica's matrix and set the running total to the sum over I <= J (the
// diagonal holds flexible molecules' internal energies).
__global__ __launch_bounds__(kWG) void prepare_finish_kernel(PairParams p)
{
    __shared__ double red[kWavesPerWG];
    const int r = blockIdx.x;
    const int M = p.n_mol;
    double *mat = p.emat + (size_t)r * M * M;
    double e = 0.0;
    for (int k = threadIdx.x; k < M * M; k += kWG) {
        const int I = k / M, J = k % M;
        if (I < J) {
            const double v = mat[k];
            mat[(size_t)J * M + I] = v;
            e += v;
        } else if (I == J) {
            e += mat[k];  // internal energy of a flexible molecule, else 0
        }
    }
    e = wave_sum(e);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kWavesPerWG; ++w) t += red[w];
        p.energy[r] = t;
        p.done[r] = 0u;
    }
    for (int J = threadIdx.x; J < M; J += kWG) {
        p.rowbuf[(size_t)r * M + J] = 0.0;
        p.rowfix[(size_t)r * M + J] = 0ull;
    }
}

struct ExchangeParams {
    int n_rep;
    uint64_t seed;
    double *temperature;
    const double *energy;
    int *label;
    int *rank_of;
    int *rep_at_rank;
    long long *counters;
    long long *round;
    unsigned long long *stats;  // tests, accepted
};

// One exchange round: src/Simulation.cpp:770-795 (pairs of adjacent
// temperature ranks starting at the round's parity, delta and min(1, e^delta)
// in FP64) and src/Replica.cpp:188-210 (temperature, label and statistics
// swap, coordinates stay).  The pairs of one round are disjoint, so lanes
// take them in parallel.
__global__ __launch_bounds__(kWave) void exchange_kernel(ExchangeParams e)
{
    const long long k = *e.round;
    const int offset = (int)(k & 1);
    const float factor = kKcal / kRgas;
    unsigned long long tests = 0, accepted = 0;
    for (int ti = offset + 2 * (int)threadIdx.x; ti < e.n_rep - 1; ti += 2 * kWave) {
        const int i = e.rep_at_rank[ti], j = e.rep_at_rank[ti + 1];
        const double delta =
            (1.0 / e.temperature[i] - 1.0 / e.temperature[j]) * (e.energy[i] - e.energy[j]) * (double)factor;
        uint32_t w[4];
        philox((uint32_t)k, (uint32_t)((unsigned long long)k >> 32), (uint32_t)ti, kTagEX, (uint32_t)e.seed,
               (uint32_t)(e.seed >> 32), w);
        const double u = unit53(w[0], w[1]);
        double pacc = exp(delta);
        if (pacc > 1.0) pacc = 1.0;
        ++tests;
        if (u < pacc) {
            const double t = e.temperature[i];
            e.temperature[i] = e.temperature[j];
            e.temperature[j] = t;
            const int l = e.label[i];
            e.label[i] = e.label[j];
            e.label[j] = l;
            for (int c = 0; c < 3; ++c) {
                const long long a = e.counters[3 * i + c];
                e.counters[3 * i + c] = e.counters[3 * j + c];
                e.counters[3 * j + c] = a;
            }
            e.rank_of[i] = ti + 1;
            e.rank_of[j] = ti;
            e.rep_at_rank[ti] = j;
            e.rep_at_rank[ti + 1] = i;
            ++accepted;
        }
    }
    if (tests) {
        atomicAdd(e.stats, tests);
        atomicAdd(e.stats + 1, accepted);
    }
    __syncthreads();
    if (threadIdx.x == 0) *e.round = k + 1;
}

// Outputs of cgppd_energy_all_noalloc: per replica the raw LJ and DH sums
// (acc[r], acc[R + r]) and their kcal/mol total, and, with mat non-null, the
// lower triangle of every replica's molecule-pair matrix mirrored from the
// upper one the full-energy kernel filled.
__global__ void noalloc_energy_finish_kernel(const double *acc, int R, int M, float rt_to_kcal, float dh_factor,
                                             double *lj, double *dh, double *total, double *mat)
{
    const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (k < (size_t)R) {
        const double l = acc[k], d = acc[R + k];
        if (lj) lj[k] = l;
        if (dh) dh[k] = d;
        // src/Potential.cpp:219
        if (total) total[k] = d * (double)dh_factor + l * (double)rt_to_kcal;
    }
    const size_t mm = (size_t)M * M;
    if (mat && k < (size_t)R * mm) {
        const size_t q = k % mm;
        const int I = (int)(q / M), J = (int)(q % M);
        if (I > J) mat[k] = mat[k - q + (size_t)J * M + I];
    }
}

__global__ void fill_zero_kernel(double *p, size_t n)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

// ---------------------------------------------------------------------------
// Surrogate: inference, gradients and Adam (mlp.h holds the tile code)
// ---------------------------------------------------------------------------

// Predictions for rows [32 g, 32 g + 32) of x (BF16 bits, n x 80).
__global__ __launch_bounds__(mlp::kThreads) void mlp_predict_kernel(const unsigned short *x, int n,
                                                                    const float *master, const unsigned short *wb,
                                                                    float *y)
{
    __shared__ mlp::Tile T;
    const int base = blockIdx.x * mlp::kRows;
    const int nv = min(mlp::kRows, n - base);
    for (int e = threadIdx.x; e < mlp::kRows * mlp::kFeatPad; e += mlp::kThreads) {
        const int k = e / mlp::kFeatPad, f = e % mlp::kFeatPad;
        T.x[e] = (k < nv && f < mlp::kFeat) ? x[(size_t)(base + k) * mlp::kFeat + f] : (unsigned short)0;
    }
    __syncthreads();
    mlp::forward(T, master, wb, nv);
    if (threadIdx.x < nv) y[base + threadIdx.x] = T.y[threadIdx.x];
}

// Gradient of 0.5 mean((y - clip(t))^2) over one minibatch, added into grad;
// loss (mean) added into loss[0].  Each workgroup takes 32 samples of the
// buffer (sx: BF16 features, st: targets): rows [32 g, 32 g + 32) when
// sequential, else drawn with replacement from Philox(step, sample, tag).
__global__ __launch_bounds__(mlp::kThreads) void mlp_grad_kernel(const unsigned short *sx, const float *st,
                                                                 int n_samples, int sequential, uint64_t seed,
                                                                 uint64_t step, float batch, const float *master,
                                                                 const unsigned short *wb, float *grad, double *loss)
{
    __shared__ mlp::Tile T;
    __shared__ float dy[mlp::kRows];
    __shared__ int idx[mlp::kRows];
    const int g = blockIdx.x;
    const int nv = sequential ? min(mlp::kRows, n_samples - g * mlp::kRows) : mlp::kRows;
    if (threadIdx.x < mlp::kRows) {
        int i = g * mlp::kRows + threadIdx.x;
        if (!sequential) {
            uint32_t w[4];
            philox((uint32_t)step, (uint32_t)(step >> 32), (uint32_t)i, kTagSG, (uint32_t)seed, (uint32_t)(seed >> 32),
                   w);
            i = mulhi_n(w[0], n_samples);
        }
        idx[threadIdx.x] = threadIdx.x < nv ? i : 0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < mlp::kRows * mlp::kFeatPad; e += mlp::kThreads) {
        const int k = e / mlp::kFeatPad, f = e % mlp::kFeatPad;
        T.x[e] = (k < nv && f < mlp::kFeat) ? sx[(size_t)idx[k] * mlp::kFeat + f] : (unsigned short)0;
    }
    __syncthreads();
    mlp::forward(T, master, wb, nv);
    if (threadIdx.x < 64) {
        double l = 0.0;
        if (threadIdx.x < nv) {
            const float t = fminf(fmaxf(st[idx[threadIdx.x]], -mlp::kTargetClip), mlp::kTargetClip);
            const float d = T.y[threadIdx.x] - t;
            dy[threadIdx.x] = d / batch;
            l = 0.5 * (double)d * (double)d / (double)batch;
        } else if (threadIdx.x < mlp::kRows) {
            dy[threadIdx.x] = 0.f;
        }
        l = wave_sum(l);
        if (threadIdx.x == 0) atomicAdd(loss, l);
    }
    __syncthreads();
    mlp::backward(T, wb, dy, grad);
}

// Adam on the FP32 masters (Kingma & Ba 2015, algorithm 1); bc1, bc2 are
// 1 - beta^t.  The gradient is cleared.  No FMA contraction, so the update
// matches an FP32 CPU restatement operation for operation.
__global__ void mlp_adam_kernel(float *master, float *m, float *v, float *grad, float lr, float bc1, float bc2)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mlp::kParams) return;
    const float g = grad[i];
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    master[i] = master[i] - lr * mh / (sqrtf(vh) + 1e-8f);
    grad[i] = 0.f;
}

// Body frame (rows e1, e2, e3) of a molecule from the relative positions
// of its two anchor residues; the identity when they coincide or degenerate.
__device__ inline void make_frame(float4 a, float4 b, bool same, float f[9])
{
    for (int k = 0; k < 9; ++k) f[k] = (k % 4 == 0) ? 1.0f : 0.0f;
    if (same) return;
    const float n1 = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    if (n1 == 0.0f) return;
    const float e1x = a.x / n1, e1y = a.y / n1, e1z = a.z / n1;
    const float d = b.x * e1x + b.y * e1y + b.z * e1z;
    const float px = b.x - d * e1x, py = b.y - d * e1y, pz = b.z - d * e1z;
    const float n2 = sqrtf(px * px + py * py + pz * pz);
    if (n2 == 0.0f) return;
    const float e2x = px / n2, e2y = py / n2, e2z = pz / n2;
    f[0] = e1x;
    f[1] = e1y;
    f[2] = e1z;
    f[3] = e2x;
    f[4] = e2y;
    f[5] = e2z;
    f[6] = e1y * e2z - e1z * e2y;
    f[7] = e1z * e2x - e1x * e2z;
    f[8] = e1x * e2y - e1y * e2x;
}

// The 80 features of the pair (I, J), I < J, stored through put(f, value).
template <typename Put>
__device__ inline void pair_feature_row(float4 ci, float4 cj, const float *fi, const float *fj, const float *hi,
                                        const float *hj, float box, float inv_box, Put put)
{
    float dx = cj.x - ci.x, dy = cj.y - ci.y, dz = cj.z - ci.z;
    dx -= box * rintf(dx * inv_box);
    dy -= box * rintf(dy * inv_box);
    dz -= box * rintf(dz * inv_box);
    const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
    const float ux = dist > 0.f ? dx / dist : 0.f, uy = dist > 0.f ? dy / dist : 0.f,
                uz = dist > 0.f ? dz / dist : 0.f;
    for (int k = 0; k < 32; ++k) {
        const float t = (dist - 3.0f * (float)k) * (1.0f / 3.0f);
        put(k, expf(-0.5f * t * t));
    }
    for (int i = 0; i < 3; ++i) {
        put(32 + i, fi[3 * i] * ux + fi[3 * i + 1] * uy + fi[3 * i + 2] * uz);
        put(35 + i, -(fj[3 * i] * ux + fj[3 * i + 1] * uy + fj[3 * i + 2] * uz));
    }
    for (int t = 0; t < 20; ++t) {
        put(38 + t, hi[t]);
        put(58 + t, hj[t]);
    }
    put(78, 0.f);
    put(79, 0.f);
}

// Pair q of a replica's M(M-1)/2 pairs in row-major (I, J) order.
__device__ inline void pair_of(int q, int M, int &I, int &J)
{
    I = 0;
    while (q >= M - 1 - I) {
        q -= M - 1 - I;
        ++I;
    }
    J = I + 1 + q;
}

// The surrogate of a pair is 0 beyond centre distance r_I + r_J + 15 A
// (r: bounding radius): such pairs are left out of training, of the
// surrogate matrix and of stage one.
constexpr float kSurrogatePad = 15.0f;

__device__ inline bool surrogate_near(float4 ci, float4 cj, float ri, float rj, float box, float inv_box)
{
    float dx = cj.x - ci.x, dy = cj.y - ci.y, dz = cj.z - ci.z;
    dx -= box * rintf(dx * inv_box);
    dy -= box * rintf(dy * inv_box);
    dz -= box * rintf(dz * inv_box);
    const float lim = ri + rj + kSurrogatePad;
    return dx * dx + dy * dy + dz * dz < lim * lim;
}

// Body frames of every molecule of every replica (current coordinates).
__global__ void frames_kernel(PairParams p)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= p.rep_count * p.n_mol) return;
    const int r = k / p.n_mol, m = k % p.n_mol;
    const int2 an = p.anchor[m];
    const float4 *rel = p.rel + (size_t)r * p.stride;
    float f[9];
    make_frame(rel[an.x], rel[an.y], an.x == an.y, f);
    for (int i = 0; i < 9; ++i) p.frames[(size_t)k * 9 + i] = f[i];
}

// Features of every pair of replicas [rep_begin, rep_begin + rep_count):
// BF16 rows into xb (+ targets from the MC matrix into t) or FP32 rows into xf.
// With slot != null only pairs with slot[k] >= 0 are written, to row slot[k].
__global__ void pair_features_kernel(PairParams p, unsigned short *xb, float *t, float *xf, const int *slot)
{
    const int P = p.n_mol * (p.n_mol - 1) / 2;
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= (long long)p.rep_count * P) return;
    const long long out = slot ? (long long)slot[k] : k;
    if (out < 0) return;
    const int rr = (int)(k / P), q = (int)(k % P), r = p.rep_begin + rr;
    int I, J;
    pair_of(q, p.n_mol, I, J);
    const size_t rm = (size_t)r * p.n_mol;
    const float *fI = p.frames + (rm + I) * 9, *fJ = p.frames + (rm + J) * 9;
    if (xb) {
        unsigned short *row = xb + (size_t)out * mlp::kFeat;
        pair_feature_row(p.center[rm + I], p.center[rm + J], fI, fJ, p.comp + I * 20, p.comp + J * 20, p.box,
                         p.inv_box, [&](int f, float v) { row[f] = mlp::f2bf(v); });
        t[out] = (float)p.emat[(rm + I) * p.n_mol + J];
    } else {
        float *row = xf + (size_t)out * mlp::kFeat;
        pair_feature_row(p.center[rm + I], p.center[rm + J], fI, fJ, p.comp + I * 20, p.comp + J * 20, p.box,
                         p.inv_box, [&](int f, float v) { row[f] = v; });
    }
}

// 1 for the pairs of replicas [rep_begin, rep_begin + rep_count) within the
// surrogate's range, else 0.
__global__ void pair_near_kernel(PairParams p, int *flag)
{
    const int P = p.n_mol * (p.n_mol - 1) / 2;
    const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (k >= (long long)p.rep_count * P) return;
    const int r = p.rep_begin + (int)(k / P);
    int I, J;
    pair_of((int)(k % P), p.n_mol, I, J);
    const size_t rm = (size_t)r * p.n_mol;
    flag[k] = surrogate_near(p.center[rm + I], p.center[rm + J], p.radius[rm + I], p.radius[rm + J], p.box, p.inv_box)
                  ? 1
                  : 0;
}

// Surrogate pair matrix of the current state: 32 pairs per workgroup over
// all replicas' pairs; smat[r][I][J] = smat[r][J][I] = S(features(I, J)).
__global__ __launch_bounds__(mlp::kThreads) void surrogate_matrix_kernel(PairParams p)
{
    __shared__ mlp::Tile T;
    __shared__ int pr[mlp::kRows], pi[mlp::kRows], pj[mlp::kRows], pn[mlp::kRows];
    const int M = p.n_mol, P = M * (M - 1) / 2;
    const long long base = (long long)blockIdx.x * mlp::kRows, total = (long long)p.rep_count * P;
    const int nv = (int)min((long long)mlp::kRows, total - base);
    if (threadIdx.x < mlp::kRows) {
        const int k = threadIdx.x;
        unsigned short *row = T.x + k * mlp::kFeatPad;
        for (int f = mlp::kFeat; f < mlp::kFeatPad; ++f) row[f] = 0;
        if (k < nv) {
            const long long g = base + k;
            const int r = (int)(g / P);
            int I, J;
            pair_of((int)(g % P), M, I, J);
            pr[k] = r;
            pi[k] = I;
            pj[k] = J;
            const size_t rm = (size_t)r * M;
            pn[k] = surrogate_near(p.center[rm + I], p.center[rm + J], p.radius[rm + I], p.radius[rm + J], p.box,
                                   p.inv_box);
            pair_feature_row(p.center[rm + I], p.center[rm + J], p.frames + (rm + I) * 9, p.frames + (rm + J) * 9,
                             p.comp + I * 20, p.comp + J * 20, p.box, p.inv_box,
                             [&](int f, float v) { row[f] = mlp::f2bf(v); });
        } else {
            for (int f = 0; f < mlp::kFeat; ++f) row[f] = 0;
        }
    }
    __syncthreads();
    mlp::forward(T, p.sp, p.swb, nv);
    if (threadIdx.x < nv) {
        const int k = threadIdx.x;
        const size_t rmm = (size_t)pr[k] * M * M;
        const float y = pn[k] ? T.y[k] : 0.0f;
        p.smat[rmm + (size_t)pi[k] * M + pj[k]] = y;
        p.smat[rmm + (size_t)pj[k] * M + pi[k]] = y;
    }
}

// Delayed acceptance, stage one for replica r (whole workgroup): the
// surrogate row of the pending proposal against the molecules within range,
// its change dS against the current surrogate matrix, and the stage-one test
// min(1, exp(-dS / kT)) with its own Philox uniform.
__device__ void stage_one(const PairParams &p, int r)
{
    __shared__ mlp::Tile T;
    __shared__ float fnew[9];
    __shared__ double red[mlp::kThreads / 64];
    __shared__ int nearJ[CGPPD_MAX_MOLECULES];
    __shared__ int n_near;
    const int M = p.n_mol, m = p.prop_mol[r];
    const float4 cm = p.prop_center[r];
    const size_t rm = (size_t)r * M;
    if (threadIdx.x == 0) {
        const int2 an = p.anchor[m];
        const int i0 = p.mol_off[m];
        const float4 *pq = p.prop_rel + (size_t)r * p.max_n;
        make_frame(pq[an.x - i0], pq[an.y - i0], an.x == an.y, fnew);
        for (int i = 0; i < 9; ++i) p.prop_frame[(size_t)r * 9 + i] = fnew[i];
        // molecules within the surrogate's range of the moved molecule
        int c = 0;
        for (int J = 0; J < M; ++J)
            if (J != m && surrogate_near(cm, p.center[rm + J], p.radius[rm + m], p.radius[rm + J], p.box, p.inv_box))
                nearJ[c++] = J;
        n_near = c;
    }
    // the current row enters with a minus sign; the new row is 0 outside the range
    double acc = 0.0;
    for (int J = threadIdx.x; J < M; J += mlp::kThreads) {
        if (J == m) continue;
        acc -= (double)p.smat[(rm + m) * M + J];
        p.srow[rm + J] = 0.0f;
    }
    __syncthreads();
    const int nn = n_near;
    for (int base = 0; base < nn; base += mlp::kRows) {
        const int nv = min(mlp::kRows, nn - base);
        if (threadIdx.x < mlp::kRows) {
            const int k = threadIdx.x;
            unsigned short *row = T.x + k * mlp::kFeatPad;
            for (int f = mlp::kFeat; f < mlp::kFeatPad; ++f) row[f] = 0;
            if (k < nv) {
                const int J = nearJ[base + k];
                auto put = [&](int f, float v) { row[f] = mlp::f2bf(v); };
                if (m < J)
                    pair_feature_row(cm, p.center[rm + J], fnew, p.frames + (rm + J) * 9, p.comp + m * 20,
                                     p.comp + J * 20, p.box, p.inv_box, put);
                else
                    pair_feature_row(p.center[rm + J], cm, p.frames + (rm + J) * 9, fnew, p.comp + J * 20,
                                     p.comp + m * 20, p.box, p.inv_box, put);
            } else {
                for (int f = 0; f < mlp::kFeat; ++f) row[f] = 0;
            }
        }
        __syncthreads();
        mlp::forward(T, p.sp, p.swb, nv);
        if (threadIdx.x < nv) {
            const int J = nearJ[base + threadIdx.x];
            const float y = T.y[threadIdx.x];
            p.srow[rm + J] = y;
            acc += (double)y;
        }
        __syncthreads();  // the next tile reuses T
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double dS = 0.0;
        for (int w = 0; w < mlp::kThreads / 64; ++w) dS += red[w];
        const uint64_t step = (uint64_t)p.steps[r];
        uint32_t w4[4];
        philox((uint32_t)step, (uint32_t)(step >> 32), (uint32_t)(r + p.rep_rng), kTagDA, (uint32_t)p.seed,
               (uint32_t)(p.seed >> 32), w4);
        const double u1 = unit53(w4[0], w4[1]);
        const double beta = 4184.0 / ((double)kRgas * p.temperature[r]);
        p.s1[r] = (dS < 0.0 || u1 < exp(-dS * beta)) ? 1 : 0;
        p.s_dS[r] = dS;
    }
}

__global__ __launch_bounds__(mlp::kThreads) void stage1_kernel(PairParams p) { stage_one(p, blockIdx.x); }

// BF16 copies of the weights for the matrix products (transposed, padded).
__global__ void mlp_pack_kernel(const float *master, unsigned short *wb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < mlp::kHidden * mlp::kFeatPad) {
        const int n = i / mlp::kFeatPad, k = i % mlp::kFeatPad;
        wb[mlp::kBW1 + i] = k < mlp::kFeat ? mlp::f2bf(master[mlp::kOW1 + k * mlp::kHidden + n]) : (unsigned short)0;
    } else if (i < mlp::kBw3) {
        const int j = i - mlp::kBW2, n = j / mlp::kHidden, k = j % mlp::kHidden;
        wb[i] = mlp::f2bf(master[mlp::kOW2 + k * mlp::kHidden + n]);
    } else if (i < mlp::kBf16Params) {
        wb[i] = mlp::f2bf(master[mlp::kOw3 + (i - mlp::kBw3)]);
    }
}

}  // namespace

// ===========================================================================
// Scratch of the allocation-free entry point
// ===========================================================================
// cgppd_energy_all_noalloc accumulates its per-replica raw LJ and DH sums in
// this file-scope device array instead of context buffers: 2 doubles per
// replica, at most CGPPD_NOALLOC_MAX_REPLICAS replicas (16 KiB per GPU,
// held until the process ends).  Every call on one GPU shares the GPU's
// copy, so calls that overlap on one GPU race.  Host code never passes the
// array's name to a kernel (in host code it is a stand-in, not a device
// address): it uses the address hipGetSymbolAddress gives for each device.
#define CGPPD_NOALLOC_SCRATCH_DOUBLES (2 * CGPPD_NOALLOC_MAX_REPLICAS)
static __device__ __align__(256) double scratch[CGPPD_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each device

// Called by cgppd_create with `device` current: looks up this device's
// address of the scratch array once.
hipError_t resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double *>(ptr);
    return hipSuccess;
}

double *scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}
}  // namespace

// ===========================================================================
// Host side
// ===========================================================================
// At most this many replica groups (streams) per context: the box's GPUs
// serve 4 hardware queues per process.
constexpr int kMaxGroups = 4;

struct cgppd_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    std::string err;

    bool have_ff = false, have_box = false, have_sys = false, have_rep = false, mc_ready = false;
    float eps[400], sigma[20], charge[20];
    float box = 0.0f;
    float rt_to_kcal = 0.0f, dh_factor = 0.0f, r0f = 0.0f;

    int M = 0, N = 0, stride = 0, n_nc = 0, max_n = 0;
    std::vector<int> mol_off, type, code;
    // device residue order: within each molecule charged residues first, then
    // by type (perm[device index] = caller's index, inv the reverse)
    std::vector<int> perm, inv, type_dev;
    std::vector<int> run_left;  // device order: residues left in the residue's type run, itself included
    int *d_run_of = nullptr, *d_run_start = nullptr;
    unsigned *d_run_code = nullptr;
    int R = 0;
    uint64_t seed = 0;

    Coef *d_table_raw = nullptr, *d_table_scaled = nullptr;
    RowCoef *d_row_table = nullptr;
    int *d_mol_off = nullptr, *d_mol_of = nullptr;
    float4 *d_pos = nullptr, *d_rel = nullptr, *d_center = nullptr;
    float *d_radius = nullptr;
    bool fast_image = true;
    float lj_reach = 0.0f;   // PairParams::lj_reach of the current table
    bool rep_crowd = false;  // repulsive crowding (cgppd_set_repulsive_crowding)
    bool rigid_only = false;  // cgppd_set_rigid_potential_only: no internal energy
    double *d_rowbuf = nullptr, *d_emat = nullptr, *d_energy = nullptr, *d_temp = nullptr;
    double *d_rcos = nullptr, *d_rsin = nullptr;
    float *d_tstep = nullptr;
    unsigned *d_done = nullptr;
    float4 *d_prop_pos = nullptr, *d_prop_rel = nullptr, *d_prop_center = nullptr;
    float *d_prop_radius = nullptr;
    int4 *d_prop_moved = nullptr;
    int *d_prop_mol = nullptr, *d_prop_nib = nullptr;
    int *d_prop_n = nullptr, *d_prop_items = nullptr;
    int4 *d_prop_hdr = nullptr;
    long long *d_counters = nullptr, *d_steps = nullptr, *d_ex_round = nullptr;
    unsigned long long *d_ex_stats = nullptr;
    unsigned long long *d_pair_evals = nullptr;
    int *d_label = nullptr, *d_rank_of = nullptr, *d_rep_at_rank = nullptr;
    double *d_out = nullptr;      // 2R: lj, dh
    double *d_out_mat = nullptr;  // R*M*M
    double *h_out = nullptr;      // pinned 2R

    // flexible molecules: per-molecule topology (n = 0: rigid), the chain
    // index of every residue (caller order), the torsion table, and the
    // device lists derived from them (rebuilt when topo_dirty)
    std::vector<cgppd_topo::Molecule> topo;
    std::vector<int> chain_index;
    double tor_v[1600] = {0.0}, tor_d[1600] = {0.0};
    bool topo_dirty = true;
    int n_flex = 0;
    std::vector<int> mol_flex_host;  // M: 1 = flexible
    int ilane_max_items = 0;          // most i-lane items one replica's proposal can have (set by mc_prepare)
    int *d_flex_mols = nullptr, *d_flex_item_off = nullptr, *d_mol_flex = nullptr;
    int4 *d_rtag = nullptr;
    int *d_bond_off = nullptr, *d_angle_off = nullptr, *d_tors_off = nullptr, *d_ind_off = nullptr;
    int2 *d_bonds = nullptr, *d_ind = nullptr;
    int4 *d_angles = nullptr, *d_tors = nullptr, *d_crank = nullptr;
    int *d_local_off = nullptr, *d_local = nullptr, *d_crank_off = nullptr;
    int *d_flexr_off = nullptr, *d_flexr = nullptr, *d_nb_off = nullptr, *d_br_off = nullptr, *d_br = nullptr;
    double *d_tor = nullptr;      // 2 x 1600: V then delta
    double *d_out_int = nullptr;  // R * 5
    double *d_out_scratch = nullptr;  // 2R: LJ/DH of internal-only queries

    // surrogate: FP32 masters, Adam moments, gradient, packed BF16 weights,
    // the sample buffer (BF16 features, FP32 targets) and a scratch batch
    bool have_sur = false;
    long long adam_t = 0;
    float *d_sp = nullptr, *d_sm = nullptr, *d_sv = nullptr, *d_sg = nullptr;
    unsigned short *d_swb = nullptr;
    double *d_sloss = nullptr;
    unsigned short *d_sx = nullptr;
    float *d_st = nullptr;
    size_t s_cap = 0, s_n = 0;
    unsigned short *d_tx = nullptr;
    float *d_tt = nullptr, *d_ty = nullptr;
    size_t t_cap = 0;
    // delayed acceptance: anchors and composition of every molecule (set with
    // the replicas), the surrogate pair matrix, frames and stage-one state
    int2 *d_anchor = nullptr;
    float *d_comp = nullptr;
    bool da = false, smat_dirty = true;
    bool s1_valid = false;  // the pending proposals' stage-one results are current
    float *d_smat = nullptr, *d_srow = nullptr, *d_frames = nullptr, *d_prop_frame = nullptr;
    int *d_s1 = nullptr;
    double *d_s_dS = nullptr;
    long long *d_s1_rej = nullptr;

    // speculative windows (rigid systems): the caller's choice (0:
    // automatic, 1: off, else slots per window), the window buffers (for
    // win_alloc slots), and the MC steps every replica has taken (they all
    // advance together; the host replays the molecule draws from it to
    // count the launches a run of steps needs)
    int window = 0;
    int win_alloc = 0;
    int4 *d_whdr = nullptr;
    float4 *d_wcen = nullptr;
    int *d_wlen = nullptr;
    long long *d_wstop = nullptr;
    float4 *d_wprop_pos = nullptr, *d_wprop_rel = nullptr;
    unsigned long long *d_wrowfix = nullptr;
    double *d_wrowbuf = nullptr;
    double *d_wu = nullptr;
    WinDraw *d_wd = nullptr;
    long long h_step = 0;

    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    int graph_groups = 0;
    int graph_win = 0;  // window slots of the captured graph (1: one step per launch pair)
    int graph_L = 0;    // window launch pairs in the captured graph

    // replica groups: the MC steps of each group run on a stream of their
    // own (group 0 on the context stream), so that one group's finish
    // kernel and launch gaps overlap another group's pair kernel
    int replica_groups = 0;  // caller's choice (0: automatic)
    hipStream_t gstream[kMaxGroups - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kMaxGroups - 1] = {};

    int wg_per_cu = 0;
    int pair_layout = CGPPD_LAYOUT_AUTO;  // caller's choice
    int tuned_layout = 0, tuned_wg = 0;    // set by cgppd_autotune
    double tune_ms[3][7] = {};             // its timings: j-lanes, j-lanes x4, i-lanes; 2..8 per CU (0: not run)
    int tune_win[3][7] = {};               // the window slots of each timing
    int tuned_window = 0;                  // the autotuner's window slots (0: not tuned)
    bool timing = false;
    std::vector<hipEvent_t> events;
    double t_mean_ms = 0.0;
    long long t_launches = 0;
};

namespace {

int fail(cgppd_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define CG_HIP(ctx, expr)                                                                               \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail((ctx), e_ == hipErrorOutOfMemory ? CGPPD_E_NOMEM : CGPPD_E_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                              \
    } while (0)

template <typename T>
int dev_alloc(cgppd_ctx *ctx, T **p, size_t n)
{
    if (*p) {
        CG_HIP(ctx, hipFree(*p));
        *p = nullptr;
    }
    if (n == 0) n = 1;
    CG_HIP(ctx, hipMalloc((void **)p, n * sizeof(T)));
    return CGPPD_OK;
}

template <typename T>
void dev_free(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void model_constants(float *rt, float *dhc, float *r0)
{
    // src/constants.h:8-26, evaluated in float as the reference does
    const float e_charge = 1.602176487e-19f, n_a = 6.02214129e23f, eps0 = 8.85418782e-12f;
    const float ang = 1e-10f, d_water = 80.0f, t_room = 298.0f;
    const float pi_f = (float)3.14159265358979323846264338327950288L;
    *rt = t_room * kRgas / kKcal;
    const float e2 = e_charge * e_charge;
    *dhc = e2 * n_a / (4.0f * pi_f * eps0 * kKcal * ang * d_water);
    *r0 = (float)std::pow(2.0, (double)(1.0f / 6.0f));
}

void destroy_graph(cgppd_ctx *ctx)
{
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_steps = 0;
    ctx->graph_win = 0;
    ctx->graph_L = 0;
}

void free_windows(cgppd_ctx *ctx)
{
    dev_free(ctx->d_whdr);
    dev_free(ctx->d_wcen);
    dev_free(ctx->d_wlen);
    dev_free(ctx->d_wstop);
    dev_free(ctx->d_wprop_pos);
    dev_free(ctx->d_wprop_rel);
    dev_free(ctx->d_wrowfix);
    dev_free(ctx->d_wrowbuf);
    dev_free(ctx->d_wu);
    dev_free(ctx->d_wd);
    ctx->win_alloc = 0;
}

void free_replicas(cgppd_ctx *ctx)
{
    destroy_graph(ctx);
    free_windows(ctx);
    ctx->h_step = 0;
    dev_free(ctx->d_pos);
    dev_free(ctx->d_rel);
    dev_free(ctx->d_center);
    dev_free(ctx->d_radius);
    dev_free(ctx->d_rowbuf);
    dev_free(ctx->d_emat);
    dev_free(ctx->d_energy);
    dev_free(ctx->d_temp);
    dev_free(ctx->d_rcos);
    dev_free(ctx->d_rsin);
    dev_free(ctx->d_tstep);
    dev_free(ctx->d_done);
    dev_free(ctx->d_prop_pos);
    dev_free(ctx->d_prop_rel);
    dev_free(ctx->d_prop_center);
    dev_free(ctx->d_prop_radius);
    dev_free(ctx->d_prop_moved);
    dev_free(ctx->d_prop_mol);
    dev_free(ctx->d_prop_nib);
    dev_free(ctx->d_prop_n);
    dev_free(ctx->d_prop_items);
    dev_free(ctx->d_prop_hdr);
    dev_free(ctx->d_counters);
    dev_free(ctx->d_steps);
    dev_free(ctx->d_ex_round);
    dev_free(ctx->d_ex_stats);
    dev_free(ctx->d_pair_evals);
    dev_free(ctx->d_label);
    dev_free(ctx->d_rank_of);
    dev_free(ctx->d_rep_at_rank);
    dev_free(ctx->d_out);
    dev_free(ctx->d_out_mat);
    dev_free(ctx->d_out_int);
    dev_free(ctx->d_out_scratch);
    dev_free(ctx->d_anchor);
    dev_free(ctx->d_comp);
    dev_free(ctx->d_smat);
    dev_free(ctx->d_srow);
    dev_free(ctx->d_frames);
    dev_free(ctx->d_prop_frame);
    dev_free(ctx->d_s1);
    dev_free(ctx->d_s_dS);
    dev_free(ctx->d_s1_rej);
    ctx->da = false;
    ctx->smat_dirty = true;
    if (ctx->h_out) (void)hipHostFree(ctx->h_out);
    ctx->h_out = nullptr;
    ctx->have_rep = false;
    ctx->mc_ready = false;
    ctx->R = 0;
}

int build_tables(cgppd_ctx *ctx)
{
    std::vector<Coef> raw(400), scaled(400);
    std::vector<RowCoef> rowt(400);
    float rc_max = 0.0f;
    for (int tj = 0; tj < 20; ++tj)
        for (int ti = 0; ti < 20; ++ti) {
            // src/Potential.cpp:14-17, float arithmetic as in the reference
            const float Eij = kLambda * (ctx->eps[ti * 20 + tj] - kE0);
            const float sij = 0.5f * (ctx->sigma[ti] + ctx->sigma[tj]);
            Coef c;
            c.sig2 = sij * sij;
            c.A = -4.0f * Eij;
            const float rc = sij * ctx->r0f;
            c.rc2 = Eij > 0.0f ? rc * rc : -1.0f;
            if (Eij > 0.0f) rc_max = std::max(rc_max, rc);
            c.qq = ctx->charge[ti] * ctx->charge[tj];
            raw[tj * 20 + ti] = c;
            Coef s = c;
            s.A = c.A * ctx->rt_to_kcal;
            s.qq = c.qq * ctx->dh_factor;
            scaled[tj * 20 + ti] = s;
            RowCoef w;
            const float s6 = c.sig2 * c.sig2 * c.sig2;
            w.B = s.A * (s6 * s6);
            w.C = s.A * s6;
            w.qq = s.qq;
            w.K = Eij > 0.0f ? -0.5f * s.A : FLT_MAX;
            rowt[tj * 20 + ti] = w;
        }
    // a pair farther apart than this is past every switch distance (with a
    // margin far above the FP32 rounding of r^2 and of rc^2)
    ctx->lj_reach = rc_max * 1.001f + 0.05f;
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->d_table_raw, 400))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_table_scaled, 400))) return rc;
    CG_HIP(ctx, hipMemcpy(ctx->d_table_raw, raw.data(), sizeof(Coef) * 400, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_table_scaled, scaled.data(), sizeof(Coef) * 400, hipMemcpyHostToDevice));
    if ((rc = dev_alloc(ctx, &ctx->d_row_table, 400))) return rc;
    CG_HIP(ctx, hipMemcpy(ctx->d_row_table, rowt.data(), sizeof(RowCoef) * 400, hipMemcpyHostToDevice));
    return CGPPD_OK;
}

uint32_t residue_code(const cgppd_ctx *ctx, int t)
{
    return ((uint32_t)t << kRowShift) | (ctx->charge[t] != 0.0f ? kChargedBit : 0u);
}

// The word of device residue i: its code and the length of the rest of its
// type run (the row kernel reads run boundaries from the records).
uint32_t residue_word(const cgppd_ctx *ctx, int i)
{
    const uint32_t left = (uint32_t)std::min<long>(ctx->run_left[i], (long)kRunMax);
    return residue_code(ctx, ctx->type_dev[i]) | left << kRunShift;
}

// The residue code travels in the w lane of the float4 coordinates.
float code_bits(uint32_t c)
{
    float f;
    std::memcpy(&f, &c, sizeof(f));
    return f;
}

// Device residue order and the constant-type runs of each molecule.  Needs
// the topology and the charges of the force field.
int build_order(cgppd_ctx *ctx)
{
    const int M = ctx->M, N = ctx->N;
    ctx->perm.resize(N);
    ctx->inv.resize(N);
    ctx->type_dev.resize(N);
    std::vector<int> run_of(N), run_start;
    std::vector<unsigned> run_code;
    for (int m = 0; m < M; ++m) {
        const int a = ctx->mol_off[m], b = ctx->mol_off[m + 1];
        for (int i = a; i < b; ++i) ctx->perm[i] = i;
        std::stable_sort(ctx->perm.begin() + a, ctx->perm.begin() + b, [&](int x, int y) {
            const bool cx = ctx->charge[ctx->type[x]] != 0.0f, cy = ctx->charge[ctx->type[y]] != 0.0f;
            if (cx != cy) return cx;
            return ctx->type[x] < ctx->type[y];
        });
        for (int i = a; i < b; ++i) {
            ctx->inv[ctx->perm[i]] = i;
            ctx->type_dev[i] = ctx->type[ctx->perm[i]];
            if (i == a || ctx->type_dev[i] != ctx->type_dev[i - 1]) {
                run_start.push_back(i);
                run_code.push_back(residue_code(ctx, ctx->type_dev[i]));
            }
            run_of[i] = (int)run_start.size() - 1;
        }
    }
    run_start.push_back(N);
    ctx->run_left.resize(N);
    for (size_t k = 0; k + 1 < run_start.size(); ++k)
        for (int i = run_start[k]; i < run_start[k + 1]; ++i) ctx->run_left[i] = run_start[k + 1] - i;
    ctx->topo_dirty = true;
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->d_run_of, N))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_run_start, run_start.size()))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_run_code, run_code.size()))) return rc;
    CG_HIP(ctx, hipMemcpy(ctx->d_run_of, run_of.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_run_start, run_start.data(), sizeof(int) * run_start.size(), hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_run_code, run_code.data(), sizeof(unsigned) * run_code.size(),
                          hipMemcpyHostToDevice));
    return CGPPD_OK;
}

// Bounding radius of a molecule: max |relative position|.
float molecule_radius(const float4 *rel, int n)
{
    double r2 = 0.0;
    for (int k = 0; k < n; ++k)
        r2 = std::max(r2, (double)rel[k].x * rel[k].x + (double)rel[k].y * rel[k].y + (double)rel[k].z * rel[k].z);
    return (float)std::sqrt(r2);
}

// Centre and relative positions from absolute coordinates, with the float
// accumulation of src/Molecule.cpp:590-609 in the caller's residue order;
// writes pos and rel (w = code) in device order for the molecule starting
// at residue i0, and returns the centre and radius.
void split_molecule(const cgppd_ctx *ctx, int i0, int n, const float *xyz, float4 *pos, float4 *rel, float4 &centre,
                    float &radius)
{
    float cx = 0.f, cy = 0.f, cz = 0.f;
    for (int k = 0; k < n; ++k) {
        cx += xyz[3 * k];
        cy += xyz[3 * k + 1];
        cz += xyz[3 * k + 2];
    }
    const float fn = (float)n;
    cx /= fn;
    cy /= fn;
    cz /= fn;
    for (int k = 0; k < n; ++k) {
        const int d = ctx->inv[i0 + k] - i0;
        const float w = code_bits(residue_word(ctx, i0 + d));
        pos[d] = make_float4(xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], w);
        rel[d] = make_float4(xyz[3 * k] - cx, xyz[3 * k + 1] - cy, xyz[3 * k + 2] - cz, w);
    }
    centre = make_float4(cx, cy, cz, 0.0f);
    radius = molecule_radius(rel, n);
}

template <typename T>
int upload_vec(cgppd_ctx *ctx, T **d, const std::vector<T> &h)
{
    int rc = dev_alloc(ctx, d, h.size());
    if (rc) return rc;
    if (!h.empty()) CG_HIP(ctx, hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return CGPPD_OK;
}

// Device lists of the flexible molecules in device residue indices: the
// exclusion tags (the reference packs them into pos.w / meta.w,
// src/Residue.cpp:41-59), bonded lists and move sets.
int sync_topology(cgppd_ctx *ctx)
{
    if (!ctx->topo_dirty) return CGPPD_OK;
    const int M = ctx->M, N = ctx->N;
    if ((int)ctx->topo.size() != M) ctx->topo.assign(M, cgppd_topo::Molecule());
    if (ctx->rep_crowd)
        for (int m = ctx->n_nc; m < M; ++m)
            if (ctx->topo[m].flexible)
                return fail(ctx, CGPPD_E_ARG, "molecule %d is a crowder: repulsive crowding needs rigid crowders", m);
    if ((int)ctx->chain_index.size() != N) ctx->chain_index.assign(N, 0);
    std::vector<int4> rtag(N);
    std::vector<int> mol_flex(M, 0), flex_mols, flex_item_off{0};
    std::vector<int> bond_off{0}, angle_off{0}, tors_off{0}, ind_off{0}, local_off{0}, crank_off{0}, flexr_off{0};
    std::vector<int> nb_off{0}, br_off{0}, local, flexr, br;
    std::vector<int2> bonds, ind;
    std::vector<int4> angles, tors, crank;
    const auto &inv = ctx->inv;
    for (int m = 0; m < M; ++m) {
        const int i0 = ctx->mol_off[m], n = ctx->mol_off[m + 1] - i0;
        const cgppd_topo::Molecule &t = ctx->topo[m];
        const bool flex = t.flexible && t.n == n;
        for (int k = 0; k < n; ++k) {
            int4 g = make_int4(0, 0, 0, k);
            if (flex) {
                g.x = k < t.vertex_count ? ctx->chain_index[i0 + k] + 1 : 0;
                g.y = t.domain_of[k] + 1;
                g.z = t.segbond_of[k] + 1;
            }
            rtag[inv[i0 + k]] = g;
        }
        if (flex) {
            mol_flex[m] = 1;
            flex_mols.push_back(m);
            const int nib = (n + kIBlock - 1) / kIBlock;
            flex_item_off.push_back(flex_item_off.back() + nib * intra_chunks(n) + 1);
            auto d = [&](int v) { return inv[i0 + v]; };
            for (const auto &b : t.bonds) bonds.push_back(make_int2(d(b[0]), d(b[1])));
            for (const auto &a : t.angles) angles.push_back(make_int4(d(a[0]), d(a[1]), d(a[2]), 0));
            for (const auto &q : t.torsions) tors.push_back(make_int4(d(q[0]), d(q[1]), d(q[2]), d(q[3])));
            for (const auto &q : t.indirect) ind.push_back(make_int2(d(q[0]), d(q[1])));
            for (int v : t.local) local.push_back(d(v));
            // crankshaft uses the residues before and after in the caller's
            // order (src/Molecule.cpp:306,318); residues at either end of the
            // molecule have no such pair and are left out
            for (int v : t.crankshaft)
                if (v > 0 && v < n - 1) crank.push_back(make_int4(d(v - 1), d(v), d(v + 1), 0));
            for (size_t e = 0; e < t.flex.size(); ++e) {
                flexr.push_back(d(t.flex[e]));
                for (int k = t.flex_nb_off[e]; k < t.flex_nb_off[e + 1]; ++k) {
                    for (int b = t.branch_off[k]; b < t.branch_off[k + 1]; ++b) br.push_back(d(t.branch[b]));
                    br_off.push_back((int)br.size());
                }
                nb_off.push_back(nb_off.back() + (t.flex_nb_off[e + 1] - t.flex_nb_off[e]));
            }
        }
        bond_off.push_back((int)bonds.size());
        angle_off.push_back((int)angles.size());
        tors_off.push_back((int)tors.size());
        ind_off.push_back((int)ind.size());
        local_off.push_back((int)local.size());
        crank_off.push_back((int)crank.size());
        flexr_off.push_back((int)flexr.size());
    }
    std::vector<double> tor(6400);
    std::copy(ctx->tor_v, ctx->tor_v + 1600, tor.begin());
    std::copy(ctx->tor_d, ctx->tor_d + 1600, tor.begin() + 1600);
    for (int k = 0; k < 1600; ++k) {
        tor[3200 + k] = std::cos(ctx->tor_d[k]);
        tor[4800 + k] = std::sin(ctx->tor_d[k]);
    }
    int rc;
    if ((rc = upload_vec(ctx, &ctx->d_rtag, rtag)) || (rc = upload_vec(ctx, &ctx->d_mol_flex, mol_flex)) ||
        (rc = upload_vec(ctx, &ctx->d_flex_mols, flex_mols)) ||
        (rc = upload_vec(ctx, &ctx->d_flex_item_off, flex_item_off)) ||
        (rc = upload_vec(ctx, &ctx->d_bond_off, bond_off)) || (rc = upload_vec(ctx, &ctx->d_bonds, bonds)) ||
        (rc = upload_vec(ctx, &ctx->d_angle_off, angle_off)) || (rc = upload_vec(ctx, &ctx->d_angles, angles)) ||
        (rc = upload_vec(ctx, &ctx->d_tors_off, tors_off)) || (rc = upload_vec(ctx, &ctx->d_tors, tors)) ||
        (rc = upload_vec(ctx, &ctx->d_ind_off, ind_off)) || (rc = upload_vec(ctx, &ctx->d_ind, ind)) ||
        (rc = upload_vec(ctx, &ctx->d_local_off, local_off)) || (rc = upload_vec(ctx, &ctx->d_local, local)) ||
        (rc = upload_vec(ctx, &ctx->d_crank_off, crank_off)) || (rc = upload_vec(ctx, &ctx->d_crank, crank)) ||
        (rc = upload_vec(ctx, &ctx->d_flexr_off, flexr_off)) || (rc = upload_vec(ctx, &ctx->d_flexr, flexr)) ||
        (rc = upload_vec(ctx, &ctx->d_nb_off, nb_off)) || (rc = upload_vec(ctx, &ctx->d_br_off, br_off)) ||
        (rc = upload_vec(ctx, &ctx->d_br, br)) || (rc = upload_vec(ctx, &ctx->d_tor, tor)))
        return rc;
    ctx->n_flex = (int)flex_mols.size();
    ctx->mol_flex_host = mol_flex;
    ctx->topo_dirty = false;
    destroy_graph(ctx);
    return CGPPD_OK;
}

int pair_layout_of(const cgppd_ctx *ctx);
int row_lanes_of(int layout);
int window_of(const cgppd_ctx *ctx);

PairParams base_params(cgppd_ctx *ctx)
{
    PairParams p;
    std::memset(&p, 0, sizeof(p));
    p.n_mol = ctx->M;
    p.n_res = ctx->N;
    p.stride = ctx->stride;
    p.n_active = ctx->M;
    p.crowd_mol = ctx->rep_crowd ? ctx->n_nc : ctx->M;
    p.ib_max = (ctx->max_n + kIBlock - 1) / kIBlock;
    p.jc_count = (ctx->N + kJChunk - 1) / kJChunk;
    p.mol_off = ctx->d_mol_off;
    p.mol_of = ctx->d_mol_of;
    p.run_of = ctx->d_run_of;
    p.run_start = ctx->d_run_start;
    p.run_code = ctx->d_run_code;
    p.rep_begin = 0;
    p.rep_count = ctx->R;
    p.rep_rng = 0;
    p.xcd_map = 0;
    p.table = ctx->d_table_raw;
    p.table_scaled = ctx->d_table_scaled;
    p.box = ctx->box;
    p.inv_box = 1.0f / ctx->box;
    p.half_box = 0.5f * ctx->box;
    p.image_reach = 0.45f * ctx->box;
    p.pos = ctx->d_pos;
    p.rel = ctx->d_rel;
    p.center = ctx->d_center;
    p.radius = ctx->d_radius;
    p.fast_image = ctx->fast_image ? 1 : 0;
    p.lj_reach = ctx->lj_reach;
    p.out_lj = ctx->d_out;
    p.out_dh = ctx->d_out + ctx->R;
    p.out_mat = nullptr;
    p.rt_to_kcal = ctx->rt_to_kcal;
    p.dh_factor = ctx->dh_factor;
    p.row_table = ctx->d_row_table;
    for (int t = 0; t < CGPPD_AA_COUNT; ++t) p.charge[t] = ctx->charge[t];
    p.seed = ctx->seed;
    p.rowbuf = ctx->d_rowbuf;
    p.rowfix = reinterpret_cast<unsigned long long *>(ctx->d_rowbuf + (size_t)ctx->R * ctx->M);
    p.done = ctx->d_done;
    p.max_n = ctx->max_n;
    p.prop_pos = ctx->d_prop_pos;
    p.prop_rel = ctx->d_prop_rel;
    p.prop_center = ctx->d_prop_center;
    p.prop_radius = ctx->d_prop_radius;
    p.prop_moved = ctx->d_prop_moved;
    p.prop_mol = ctx->d_prop_mol;
    p.prop_nib = ctx->d_prop_nib;
    p.prop_n = ctx->d_prop_n;
    p.prop_items = ctx->d_prop_items;
    p.prop_hdr = ctx->d_prop_hdr;
    p.row_j = kWave * row_lanes_of(pair_layout_of(ctx));
    p.win_log2 = 0;
    for (int w = window_of(ctx); w > 1; w >>= 1) ++p.win_log2;
    p.whdr = ctx->d_whdr;
    p.wcen = ctx->d_wcen;
    p.wlen = ctx->d_wlen;
    p.wstop = ctx->d_wstop;
    p.wprop_pos = ctx->d_wprop_pos;
    p.wprop_rel = ctx->d_wprop_rel;
    p.wrowfix = ctx->d_wrowfix;
    p.wrowbuf = ctx->d_wrowbuf;
    p.wu = ctx->d_wu;
    p.wd = ctx->d_wd;
    p.emat = ctx->d_emat;
    p.energy = ctx->d_energy;
    p.temperature = ctx->d_temp;
    p.tstep = ctx->d_tstep;
    p.rcos = ctx->d_rcos;
    p.rsin = ctx->d_rsin;
    p.counters = ctx->d_counters;
    p.steps = ctx->d_steps;
    p.pair_evals = ctx->d_pair_evals;
    p.n_flex = ctx->n_flex;
    p.internal = ctx->rigid_only ? 0 : 1;
    p.flex_mols = ctx->d_flex_mols;
    p.flex_item_off = ctx->d_flex_item_off;
    p.mol_flex = ctx->d_mol_flex;
    p.rtag = ctx->d_rtag;
    p.bond_off = ctx->d_bond_off;
    p.bonds = ctx->d_bonds;
    p.angle_off = ctx->d_angle_off;
    p.angles = ctx->d_angles;
    p.tors_off = ctx->d_tors_off;
    p.tors = ctx->d_tors;
    p.ind_off = ctx->d_ind_off;
    p.ind = ctx->d_ind;
    p.local_off = ctx->d_local_off;
    p.local = ctx->d_local;
    p.crank_off = ctx->d_crank_off;
    p.crank = ctx->d_crank;
    p.flexr_off = ctx->d_flexr_off;
    p.flexr = ctx->d_flexr;
    p.nb_off = ctx->d_nb_off;
    p.br_off = ctx->d_br_off;
    p.br = ctx->d_br;
    p.tor_v = ctx->d_tor;
    p.tor_d = ctx->d_tor ? ctx->d_tor + 1600 : nullptr;
    p.tor_cd = ctx->d_tor ? ctx->d_tor + 3200 : nullptr;
    p.tor_sd = ctx->d_tor ? ctx->d_tor + 4800 : nullptr;
    p.out_int = ctx->d_out_int;
    p.da = ctx->da ? 1 : 0;
    p.s1 = ctx->d_s1;
    p.s_dS = ctx->d_s_dS;
    p.smat = ctx->d_smat;
    p.srow = ctx->d_srow;
    p.frames = ctx->d_frames;
    p.prop_frame = ctx->d_prop_frame;
    p.anchor = ctx->d_anchor;
    p.comp = ctx->d_comp;
    p.sp = ctx->d_sp;
    p.swb = ctx->d_swb;
    p.s1_rej = ctx->d_s1_rej;
    return p;
}

// Body frames of every molecule of every replica into d_frames.
int launch_frames(cgppd_ctx *ctx)
{
    PairParams p = base_params(ctx);
    const int n = ctx->R * ctx->M;
    hipLaunchKernelGGL(frames_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, p);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

// Frames and the surrogate pair matrix of the current state.
int refresh_surrogate_state(cgppd_ctx *ctx)
{
    int rc = launch_frames(ctx);
    if (rc) return rc;
    const long long pairs = (long long)ctx->R * ctx->M * (ctx->M - 1) / 2;
    CG_HIP(ctx, hipMemsetAsync(ctx->d_smat, 0, sizeof(float) * (size_t)ctx->R * ctx->M * ctx->M, ctx->stream));
    if (pairs > 0) {
        PairParams p = base_params(ctx);
        hipLaunchKernelGGL(surrogate_matrix_kernel, dim3((unsigned)((pairs + mlp::kRows - 1) / mlp::kRows)),
                           dim3(mlp::kThreads), 0, ctx->stream, p);
        CG_HIP(ctx, hipGetLastError());
    }
    ctx->smat_dirty = false;
    ctx->s1_valid = false;
    return CGPPD_OK;
}

// Anchor residues (device indices) and residue-type fractions of every
// molecule, from the relative positions rel0 (caller order, one replica):
// anchor a is the residue farthest from the centre (first of equals), b the
// one farthest from the line through a; molecules of fewer than 3 residues
// use their first residue twice.
int set_anchors(cgppd_ctx *ctx, const float *rel0)
{
    const int M = ctx->M;
    std::vector<int2> an(M);
    std::vector<float> comp((size_t)M * 20, 0.0f);
    for (int m = 0; m < M; ++m) {
        const int s0 = ctx->mol_off[m], n = ctx->mol_off[m + 1] - s0;
        const float *x = rel0 + (size_t)s0 * 3;
        int a = 0, b = 0;
        if (n >= 3) {
            double best = -1.0;
            for (int k = 0; k < n; ++k) {
                const double v[3] = {x[3 * k], x[3 * k + 1], x[3 * k + 2]};
                const double d = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                if (d > best) {
                    best = d;
                    a = k;
                }
            }
            const double ax = x[3 * a], ay = x[3 * a + 1], az = x[3 * a + 2];
            const double na = std::max(std::sqrt(ax * ax + ay * ay + az * az), 1e-30);
            const double e[3] = {ax / na, ay / na, az / na};
            best = -1.0;
            for (int k = 0; k < n; ++k) {
                const double v[3] = {x[3 * k], x[3 * k + 1], x[3 * k + 2]};
                const double d = v[0] * e[0] + v[1] * e[1] + v[2] * e[2];
                const double px = v[0] - d * e[0], py = v[1] - d * e[1], pz = v[2] - d * e[2];
                const double q = px * px + py * py + pz * pz;
                if (q > best) {
                    best = q;
                    b = k;
                }
            }
        }
        an[m] = make_int2(ctx->inv[s0 + a], ctx->inv[s0 + b]);
        for (int k = 0; k < n; ++k) comp[(size_t)m * 20 + ctx->type[s0 + k]] += 1.0f;
        for (int t = 0; t < 20; ++t) comp[(size_t)m * 20 + t] = (float)((double)comp[(size_t)m * 20 + t] / (double)n);
    }
    int rc;
    if ((rc = upload_vec(ctx, &ctx->d_anchor, an))) return rc;
    return upload_vec(ctx, &ctx->d_comp, comp);
}

int grid_for(const cgppd_ctx *ctx, long long items, bool xcd)
{
    const int per_cu = ctx->wg_per_cu > 0 ? ctx->wg_per_cu : 4;
    long long g = (long long)ctx->n_cu * per_cu;
    if (xcd) {
        // each XCD takes items / 8
        long long per = (items + 7) / 8;
        g = std::min(g, per * 8);
        g = std::max(8LL, g / 8 * 8);
    } else {
        g = std::min(g, items);
        g = std::max(1LL, g);
    }
    return (int)g;
}

// The first crowder molecule whose stored topology is flexible, or -1.
int flexible_crowder(const cgppd_ctx *ctx)
{
    for (int m = ctx->n_nc; m < (int)ctx->topo.size(); ++m)
        if (ctx->topo[m].flexible) return m;
    return -1;
}

int check_ready(cgppd_ctx *ctx, bool need_rep)
{
    if (!ctx) return CGPPD_E_ARG;
    if (!ctx->have_ff) return fail(ctx, CGPPD_E_STATE, "force field not set");
    if (!ctx->have_box) return fail(ctx, CGPPD_E_STATE, "box not set");
    if (!ctx->have_sys) return fail(ctx, CGPPD_E_STATE, "system not set");
    if (need_rep && !ctx->have_rep) return fail(ctx, CGPPD_E_STATE, "replicas not set");
    if (ctx->topo_dirty) {
        CG_HIP(ctx, hipSetDevice(ctx->device));
        return sync_topology(ctx);
    }
    return CGPPD_OK;
}

// out_lj / out_dh: where the raw sums go (unscaled launches; null: the
// context's output buffer)
int launch_full(cgppd_ctx *ctx, int r_begin, int r_count, int n_active, bool scaled, double *mat,
                double *out_lj = nullptr, double *out_dh = nullptr)
{
    PairParams p = base_params(ctx);
    p.rep_begin = r_begin;
    p.rep_count = r_count;
    p.n_active = n_active;
    p.out_mat = mat;
    if (out_lj) p.out_lj = out_lj;
    if (out_dh) p.out_dh = out_dh;
    p.table = scaled ? ctx->d_table_scaled : ctx->d_table_raw;
    p.xcd_map = (r_begin == 0 && r_count % 8 == 0) ? 1 : 0;
    const long long items = (long long)r_count * n_active * p.ib_max * p.jc_count;
    if (items == 0) return CGPPD_OK;
    const int g = grid_for(ctx, items, p.xcd_map);
    if (scaled)
        hipLaunchKernelGGL(full_energy_kernel<true>, dim3(g), dim3(kWG), 0, ctx->stream, p);
    else
        hipLaunchKernelGGL(full_energy_kernel<false>, dim3(g), dim3(kWG), 0, ctx->stream, p);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

// Intra-molecular pairs and bonded terms of the flexible molecules of
// replicas [r_begin, r_begin + r_count); out_lj/out_dh receive the raw
// intra-molecular LJ/DH (unscaled launches).
int launch_intra(cgppd_ctx *ctx, int r_begin, int r_count, bool scaled, double *mat, double *out_lj,
                 double *out_dh)
{
    if (ctx->n_flex == 0 || ctx->rigid_only) return CGPPD_OK;
    PairParams p = base_params(ctx);
    p.rep_begin = r_begin;
    p.rep_count = r_count;
    p.out_mat = mat;
    p.out_lj = out_lj;
    p.out_dh = out_dh;
    p.table = scaled ? ctx->d_table_scaled : ctx->d_table_raw;
    p.xcd_map = 0;
    int per_rep = 0;
    for (int m = 0; m < ctx->M; ++m)
        if (ctx->topo[m].flexible) {
            const int n = ctx->mol_off[m + 1] - ctx->mol_off[m];
            per_rep += ((n + kIBlock - 1) / kIBlock) * ((n + kJChunk - 1) / kJChunk) + 1;
        }
    const int g = grid_for(ctx, (long long)per_rep * r_count, false);
    if (scaled)
        hipLaunchKernelGGL(intra_energy_kernel<true>, dim3(g), dim3(kWG), 0, ctx->stream, p);
    else
        hipLaunchKernelGGL(intra_energy_kernel<false>, dim3(g), dim3(kWG), 0, ctx->stream, p);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

int zero_dev(cgppd_ctx *ctx, double *p, size_t n)
{
    if (n == 0) return CGPPD_OK;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, ctx->stream, p, n);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

// Residues per lane of the row kernel in a layout (2 outside the j-lane layouts).
int row_lanes_of(int layout) { return layout == CGPPD_LAYOUT_JLANES4 ? 4 : 2; }

// Layout of the pair kernel: the caller's choice, else the autotuned one,
// else j-lanes (flexible systems always run the i-lane kernel).
int pair_layout_of(const cgppd_ctx *ctx)
{
    if (ctx->n_flex) return CGPPD_LAYOUT_ILANES;
    // the row kernel counts one replica's units in 32 bits
    if ((long long)((ctx->N + kWave - 1) / kWave) * ctx->max_n >= (1ll << 32)) return CGPPD_LAYOUT_ILANES;
    if (ctx->pair_layout == CGPPD_LAYOUT_RESIDENT) return CGPPD_LAYOUT_JLANES;  // (its pair-kernel steps)
    if (ctx->pair_layout != CGPPD_LAYOUT_AUTO) return ctx->pair_layout;
    return ctx->tuned_layout ? ctx->tuned_layout : CGPPD_LAYOUT_JLANES;
}

// Workgroups per CU: the caller's choice, else the autotuned one (when it
// was tuned for this layout), else 6 for i-lanes (measured best of 4..8 on
// a 40-chain system) and 8 for j-lanes (the occupancy limit).
int pair_wg_per_cu(const cgppd_ctx *ctx, int layout)
{
    if (ctx->wg_per_cu > 0) return ctx->wg_per_cu;
    if (ctx->tuned_wg > 0 && ctx->tuned_layout == layout) return ctx->tuned_wg;
    return layout == CGPPD_LAYOUT_ILANES ? 6 : 8;
}

// ev_start/ev_stop (optional) are recorded by the dispatch itself, so their
// elapsed time is the kernel's execution time without separate event packets.
// Dynamic LDS of the row kernel for V proposals (replicas, or window
// slots): V + 1 unit offsets (padded to 16 bytes), then, for V <= kHdrLds,
// V proposal headers and V centres.
size_t row_lds_bytes(int V)
{
    return sizeof(long long) * (size_t)((V + 2) & ~1) +
           (V <= kHdrLds ? (size_t)V * (sizeof(int4) + sizeof(float4)) : 0);
}

// Number of replica groups the MC steps use: the caller's choice, else 1;
// never more groups than replicas.
int groups_of(const cgppd_ctx *ctx)
{
    const int g = ctx->replica_groups > 0 ? ctx->replica_groups : 1;
    return std::max(1, std::min(std::min(g, kMaxGroups), ctx->R));
}

// Kernel parameters of replica group g of G: the replicas
// [g R / G, (g + 1) R / G), seen by the MC kernels as 0..n-1 (every
// per-replica pointer starts at the group's first replica).
PairParams group_params(cgppd_ctx *ctx, int g, int G)
{
    PairParams p = base_params(ctx);
    if (G <= 1) return p;
    const int R = ctx->R;
    const int r0 = (int)((long long)g * R / G), r1 = (int)((long long)(g + 1) * R / G);
    const size_t M = ctx->M, S = ctx->stride, mx = ctx->max_n, q = r0;
    p.rep_count = r1 - r0;
    p.rep_rng = r0;
    p.pos += q * S;
    p.rel += q * S;
    p.center += q * M;
    p.radius += q * M;
    p.rowbuf += q * M;
    p.rowfix += q * M;
    p.done += q;
    p.emat += q * M * M;
    p.energy += q;
    p.temperature += q;
    p.tstep += q;
    p.rcos += q;
    p.rsin += q;
    p.counters += 3 * q;
    p.steps += q;
    p.prop_pos += q * mx;
    p.prop_rel += q * mx;
    p.prop_center += q;
    p.prop_radius += q;
    p.prop_moved += q;
    p.prop_mol += q;
    p.prop_nib += q;
    p.prop_n += q;
    p.prop_items += q;
    p.prop_hdr += q;
    if (p.s1) p.s1 += q;
    if (p.s_dS) p.s_dS += q;
    if (p.smat) p.smat += q * M * M;
    if (p.srow) p.srow += q * M;
    if (p.frames) p.frames += q * M * 9;
    if (p.prop_frame) p.prop_frame += q * 9;
    if (p.s1_rej) p.s1_rej += q;
    p.out_lj = p.out_dh = p.out_mat = p.out_int = nullptr;
    return p;
}

hipStream_t group_stream(const cgppd_ctx *ctx, int g) { return g == 0 ? ctx->stream : ctx->gstream[g - 1]; }

// Streams and fork/join events of G groups (created once, default priority).
int ensure_group_streams(cgppd_ctx *ctx, int G)
{
    if (G > 1 && !ctx->ev_fork) CG_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    for (int g = 1; g < G; ++g) {
        if (!ctx->gstream[g - 1]) CG_HIP(ctx, hipStreamCreateWithFlags(&ctx->gstream[g - 1], hipStreamNonBlocking));
        if (!ctx->ev_join[g - 1]) CG_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_join[g - 1], hipEventDisableTiming));
    }
    return CGPPD_OK;
}

int launch_pair_p(cgppd_ctx *ctx, PairParams p, hipStream_t stream, hipEvent_t ev_start = nullptr,
                  hipEvent_t ev_stop = nullptr)
{
    p.table = ctx->d_table_scaled;
    p.xcd_map = 1;
    const int layout = pair_layout_of(ctx);
    int g = std::max(8, ctx->n_cu * pair_wg_per_cu(ctx, layout) / 8 * 8);
    if (ctx->n_flex) {
        // no more workgroups than the replicas' proposals can have items: a
        // workgroup without one only delays the dispatch of the others
        const long long most = (long long)ctx->ilane_max_items * p.rep_count;
        g = (int)std::max(8LL, std::min((long long)g, (most + 7) / 8 * 8));
        if (ev_start)
            hipExtLaunchKernelGGL(mc_pair_kernel<true>, dim3(g), dim3(kWG), 0, stream, ev_start, ev_stop, 0, p);
        else
            hipLaunchKernelGGL(mc_pair_kernel<true>, dim3(g), dim3(kWG), 0, stream, p);
    } else if (layout == CGPPD_LAYOUT_ILANES) {
        if (ev_start)
            hipExtLaunchKernelGGL(mc_pair_kernel<false>, dim3(g), dim3(kWG), 0, stream, ev_start, ev_stop, 0, p);
        else
            hipLaunchKernelGGL(mc_pair_kernel<false>, dim3(g), dim3(kWG), 0, stream, p);
    } else {
        const bool crowd = p.crowd_mol < p.n_mol;
        const bool win = p.win_log2 > 0;
        auto kern = layout == CGPPD_LAYOUT_JLANES4
                        ? (win ? mc_row_kernel<4, false, true> : crowd ? mc_row_kernel<4, true, false> : mc_row_kernel<4, false, false>)
                        : (win ? mc_row_kernel<2, false, true> : crowd ? mc_row_kernel<2, true, false> : mc_row_kernel<2, false, false>);
        const size_t lds = row_lds_bytes(p.rep_count << p.win_log2);
        if (ev_start)
            hipExtLaunchKernelGGL(kern, dim3(g), dim3(kWG), lds, stream, ev_start, ev_stop, 0, p);
        else
            hipLaunchKernelGGL(kern, dim3(g), dim3(kWG), lds, stream, p);
    }
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

int launch_pair(cgppd_ctx *ctx, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr)
{
    return launch_pair_p(ctx, base_params(ctx), ctx->stream, ev_start, ev_stop);
}

int launch_finish_p(cgppd_ctx *ctx, const PairParams &p, hipStream_t stream, int propose_only)
{
    if (!propose_only && !ctx->n_flex && !ctx->da)
        hipLaunchKernelGGL(mc_finish_rigid_kernel, dim3(p.rep_count), dim3(kWG), 0, stream, p);
    else
        hipLaunchKernelGGL(mc_finish_kernel, dim3(p.rep_count), dim3(kWG), 0, stream, p, propose_only);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

int launch_finish(cgppd_ctx *ctx, int propose_only)
{
    return launch_finish_p(ctx, base_params(ctx), ctx->stream, propose_only);
}

int launch_stage1(cgppd_ctx *ctx)
{
    if (!ctx->da) return CGPPD_OK;
    PairParams p = base_params(ctx);
    hipLaunchKernelGGL(stage1_kernel, dim3(ctx->R), dim3(mlp::kThreads), 0, ctx->stream, p);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

// Before MC steps with delayed acceptance: the surrogate matrix and the
// stage-one results of the pending proposals (the finish kernel computes
// them for every later proposal).
int ensure_stage1(cgppd_ctx *ctx)
{
    if (!ctx->da) return CGPPD_OK;
    int rc;
    if (ctx->smat_dirty && (rc = refresh_surrogate_state(ctx))) return rc;
    if (!ctx->s1_valid) {
        if ((rc = launch_stage1(ctx))) return rc;
        ctx->s1_valid = true;
    }
    return CGPPD_OK;
}

int launch_mc_step(cgppd_ctx *ctx)
{
    int rc = launch_pair(ctx);
    if (rc) return rc;
    return launch_finish(ctx, 0);
}

// Resident MC steps: chosen with CGPPD_LAYOUT_RESIDENT, or by
// CGPPD_LAYOUT_AUTO for small systems (at most 16384 residue pairs per
// replica step on average over the molecules: there a step is a chain of
// latencies that one workgroup per replica shortens), for systems of rigid
// molecules without delayed acceptance or repulsive crowding whose replica
// fits the kernel's LDS, with one replica group; otherwise the pair and
// finish kernels run.
bool resident_small(const cgppd_ctx *ctx)
{
    double pairs = 0.0;
    for (int m = 0; m < ctx->M; ++m) {
        const double n = ctx->mol_off[m + 1] - ctx->mol_off[m];
        pairs += n * (ctx->N - n) / ctx->M;
    }
    return pairs <= 16384.0;
}

bool resident_active(const cgppd_ctx *ctx)
{
    const bool chosen = ctx->pair_layout == CGPPD_LAYOUT_RESIDENT ||
                        (ctx->pair_layout == CGPPD_LAYOUT_AUTO && resident_small(ctx));
    return chosen && !ctx->n_flex && !ctx->da && !ctx->rep_crowd && ctx->N <= kResidentRes && ctx->M <= kResidentMol &&
           groups_of(ctx) == 1 && resident_lds_bytes(ctx->N, ctx->M, ctx->max_n) + kTableBytes + 1024 <= 65536;
}

// Slots per speculative window: windows serve rigid systems on the j-lane
// row kernel without delayed acceptance, repulsive crowding or replica
// groups (1 elsewhere: one MC step per launch pair); the caller's choice,
// else the autotuner's, else kAutoWindow for systems of at least
// kAutoWindowMol molecules (with fewer, a window ends on a repeated
// molecule after about two steps, which does not repay the decision
// kernel).
constexpr int kAutoWindow = 4;
constexpr int kAutoWindowMol = 8;
int window_of(const cgppd_ctx *ctx)
{
    if (ctx->window == 1 || ctx->n_flex || ctx->da || ctx->rep_crowd || groups_of(ctx) != 1 || resident_active(ctx))
        return 1;
    const int layout = pair_layout_of(ctx);
    if (layout != CGPPD_LAYOUT_JLANES && layout != CGPPD_LAYOUT_JLANES4) return 1;
    if (ctx->window > 1) return ctx->window;
    // the autotuner's choice holds for the layout it was tuned with
    if (ctx->tuned_window > 0 && ctx->tuned_layout == layout) return ctx->tuned_window;
    return ctx->M >= kAutoWindowMol ? kAutoWindow : 1;
}

// Window buffers for W slots (allocated and cleared outside any capture).
int ensure_windows(cgppd_ctx *ctx, int W)
{
    if (W <= 1 || ctx->win_alloc == W) return CGPPD_OK;
    destroy_graph(ctx);
    free_windows(ctx);
    const size_t R = ctx->R, M = ctx->M;
    const size_t rows = R * W * (M + W);
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->d_whdr, R * W))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_wcen, R * W))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_wlen, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_wstop, R))) return rc;
    // 8 records of padding: the row kernel reads i records in groups
    if ((rc = dev_alloc(ctx, &ctx->d_wprop_pos, R * (W - 1) * ctx->max_n + 8))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_wprop_rel, R * (W - 1) * ctx->max_n))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_wrowfix, rows))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_wrowbuf, rows))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_wu, R * W))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_wd, R * W))) return rc;
    CG_HIP(ctx, hipMemset(ctx->d_wrowfix, 0, sizeof(unsigned long long) * rows));
    CG_HIP(ctx, hipMemset(ctx->d_wrowbuf, 0, sizeof(double) * rows));
    CG_HIP(ctx, hipMemset(ctx->d_wlen, 0, sizeof(int) * R));
    CG_HIP(ctx, hipMemset(ctx->d_wprop_pos, 0, sizeof(float4) * (R * (W - 1) * ctx->max_n + 8)));
    ctx->win_alloc = W;
    return CGPPD_OK;
}

// Windows replica r needs for the steps [step0, step0 + n): the molecule
// draws replayed on the host (draw_mol), cut as window_len cuts them.
long long windows_of_replica(const cgppd_ctx *ctx, int r, long long step0, int n, int W)
{
    const long long stop = step0 + n;
    long long s = step0, count = 0;
    while (s < stop) {
        const long long at = s;
        s += window_len([&](int k) { return draw_mol(ctx->seed, r, at + k, ctx->M); }, at, stop, W);
        ++count;
    }
    return count;
}

// Launch pairs a run of n steps from step0 needs: the most windows of any
// replica.
int windows_needed(const cgppd_ctx *ctx, long long step0, int n, int W)
{
    long long most = 0;
    for (int r = 0; r < ctx->R; ++r) most = std::max(most, windows_of_replica(ctx, r, step0, n, W));
    return (int)most;
}

int launch_window_begin(cgppd_ctx *ctx, int n_steps)
{
    const PairParams p = base_params(ctx);
    hipLaunchKernelGGL(mc_window_kernel, dim3(ctx->R), dim3(kWinWG), 0, ctx->stream, p, n_steps);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

// One window launch pair: the rows of every replica's window, then the
// decisions and the next windows.
int launch_window_pair(cgppd_ctx *ctx, hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr)
{
    int rc = launch_pair(ctx, ev_start, ev_stop);
    if (rc) return rc;
    const PairParams p = base_params(ctx);
    hipLaunchKernelGGL(mc_window_kernel, dim3(ctx->R), dim3(kWinWG), 0, ctx->stream, p, -1);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

// n_steps MC steps in windows from ctx->h_step (which the caller advances),
// as runs of at most kWindowChunk steps: the host counts the next run's
// windows while the device works through the previous one (a run ends
// windows at its end, which leaves the chain unchanged).
constexpr int kWindowChunk = 8192;
int launch_window_steps(cgppd_ctx *ctx, int n_steps)
{
    const int W = window_of(ctx);
    int rc = CGPPD_OK;
    for (int done = 0; done < n_steps && !rc;) {
        const int n = std::min(n_steps - done, kWindowChunk);
        rc = launch_window_begin(ctx, n);
        // (the host replays the draws while the start kernel runs)
        const int L = rc ? 0 : windows_needed(ctx, ctx->h_step + done, n, W);
        for (int k = 0; k < L && !rc; ++k) rc = launch_window_pair(ctx);
        done += n;
    }
    return rc;
}

// n_steps resident MC steps, at most kResidentChunk per launch.
constexpr int kResidentChunk = 16384;
int launch_resident_steps(cgppd_ctx *ctx, int n_steps)
{
    PairParams p = base_params(ctx);
    p.table = ctx->d_table_scaled;
    const size_t lds = resident_lds_bytes(ctx->N, ctx->M, ctx->max_n);
    for (int done = 0; done < n_steps;) {
        ResidentParams rp;
        rp.steps = std::min(n_steps - done, kResidentChunk);
        hipLaunchKernelGGL(mc_resident_kernel, dim3(ctx->R), dim3(kWG), lds, ctx->stream, p, rp);
        CG_HIP(ctx, hipGetLastError());
        done += rp.steps;
    }
    return CGPPD_OK;
}

// n_steps MC steps of every replica.  With G > 1 replica groups the context
// stream forks into the groups' streams and joins them again, so the work
// after the call (an exchange, a read-back) still follows every group's
// steps; the same calls are recorded into a graph under stream capture.
int launch_mc_steps(cgppd_ctx *ctx, int n_steps)
{
    const int G = groups_of(ctx);
    int rc;
    if (G == 1) {
        if (resident_active(ctx)) return n_steps > 0 ? launch_resident_steps(ctx, n_steps) : CGPPD_OK;
        if (window_of(ctx) > 1) return launch_window_steps(ctx, n_steps);
        for (int k = 0; k < n_steps; ++k)
            if ((rc = launch_mc_step(ctx))) return rc;
        return CGPPD_OK;
    }
    if ((rc = ensure_group_streams(ctx, G))) return rc;
    CG_HIP(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
    for (int g = 1; g < G; ++g) CG_HIP(ctx, hipStreamWaitEvent(ctx->gstream[g - 1], ctx->ev_fork, 0));
    for (int g = 0; g < G; ++g) {
        const PairParams p = group_params(ctx, g, G);
        const hipStream_t st = group_stream(ctx, g);
        for (int k = 0; k < n_steps; ++k) {
            if ((rc = launch_pair_p(ctx, p, st))) return rc;
            if ((rc = launch_finish_p(ctx, p, st, 0))) return rc;
        }
    }
    for (int g = 1; g < G; ++g) {
        CG_HIP(ctx, hipEventRecord(ctx->ev_join[g - 1], ctx->gstream[g - 1]));
        CG_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join[g - 1], 0));
    }
    return CGPPD_OK;
}

int launch_exchange(cgppd_ctx *ctx)
{
    ExchangeParams e;
    e.n_rep = ctx->R;
    e.seed = ctx->seed;
    e.temperature = ctx->d_temp;
    e.energy = ctx->d_energy;
    e.label = ctx->d_label;
    e.rank_of = ctx->d_rank_of;
    e.rep_at_rank = ctx->d_rep_at_rank;
    e.counters = ctx->d_counters;
    e.round = ctx->d_ex_round;
    e.stats = ctx->d_ex_stats;
    hipLaunchKernelGGL(exchange_kernel, dim3(1), dim3(kWave), 0, ctx->stream, e);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

int energy_outputs(cgppd_ctx *ctx, int r, double *lj, double *dh, double *total)
{
    CG_HIP(ctx, hipMemcpyAsync(ctx->h_out, ctx->d_out, sizeof(double) * 2 * ctx->R, hipMemcpyDeviceToHost,
                               ctx->stream));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double l = ctx->h_out[r], d = ctx->h_out[ctx->R + r];
    if (lj) *lj = l;
    if (dh) *dh = d;
    // src/Potential.cpp:219
    if (total) *total = d * (double)ctx->dh_factor + l * (double)ctx->rt_to_kcal;
    return CGPPD_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int cgppd_device_count(int *count)
{
    if (!count) return CGPPD_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return CGPPD_OK;
}

int cgppd_create(cgppd_ctx **out, int device)
{
    if (!out) return CGPPD_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CGPPD_E_HIP;
    if (device < 0 || device >= n) return CGPPD_E_ARG;
    cgppd_ctx *ctx = new cgppd_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CGPPD_E_HIP;
    }
    // the allocation-free entry point's scratch address on this device
    if (resolve_scratch(device) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return CGPPD_E_HIP;
    }
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0)
        ctx->n_cu = cu;
    model_constants(&ctx->rt_to_kcal, &ctx->dh_factor, &ctx->r0f);
    *out = ctx;
    return CGPPD_OK;
}

static void free_surrogate(cgppd_ctx *ctx)
{
    dev_free(ctx->d_sp);
    dev_free(ctx->d_sm);
    dev_free(ctx->d_sv);
    dev_free(ctx->d_sg);
    dev_free(ctx->d_swb);
    dev_free(ctx->d_sloss);
    dev_free(ctx->d_sx);
    dev_free(ctx->d_st);
    dev_free(ctx->d_tx);
    dev_free(ctx->d_tt);
    dev_free(ctx->d_ty);
    ctx->s_cap = ctx->s_n = ctx->t_cap = 0;
    ctx->have_sur = false;
}

int cgppd_destroy(cgppd_ctx *ctx)
{
    if (!ctx) return CGPPD_E_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_replicas(ctx);
    dev_free(ctx->d_table_raw);
    dev_free(ctx->d_row_table);
    dev_free(ctx->d_table_scaled);
    dev_free(ctx->d_mol_off);
    dev_free(ctx->d_mol_of);
    dev_free(ctx->d_run_of);
    dev_free(ctx->d_run_start);
    dev_free(ctx->d_run_code);
    for (int **q : {&ctx->d_flex_mols, &ctx->d_flex_item_off, &ctx->d_mol_flex, &ctx->d_bond_off, &ctx->d_angle_off,
                    &ctx->d_tors_off, &ctx->d_ind_off, &ctx->d_local_off, &ctx->d_local, &ctx->d_crank_off,
                    &ctx->d_flexr_off, &ctx->d_flexr, &ctx->d_nb_off, &ctx->d_br_off, &ctx->d_br})
        dev_free(*q);
    dev_free(ctx->d_rtag);
    dev_free(ctx->d_bonds);
    dev_free(ctx->d_ind);
    dev_free(ctx->d_angles);
    dev_free(ctx->d_tors);
    dev_free(ctx->d_crank);
    dev_free(ctx->d_tor);
    free_surrogate(ctx);
    for (hipEvent_t ev : ctx->events) (void)hipEventDestroy(ev);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (int g = 0; g + 1 < kMaxGroups; ++g) {
        if (ctx->ev_join[g]) (void)hipEventDestroy(ctx->ev_join[g]);
        if (ctx->gstream[g]) (void)hipStreamDestroy(ctx->gstream[g]);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CGPPD_OK;
}

const char *cgppd_last_error(const cgppd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cgppd_model_constants(float *rt_to_kcal, float *dh_factor, float *r0_factor)
{
    float a, b, c;
    model_constants(&a, &b, &c);
    if (rt_to_kcal) *rt_to_kcal = a;
    if (dh_factor) *dh_factor = b;
    if (r0_factor) *r0_factor = c;
    return CGPPD_OK;
}

int cgppd_set_force_field(cgppd_ctx *ctx, const float *eps, const float *sigma, const float *charge)
{
    if (!ctx) return CGPPD_E_ARG;
    if (!eps || !sigma || !charge) return fail(ctx, CGPPD_E_ARG, "null force-field array");
    std::memcpy(ctx->eps, eps, sizeof(ctx->eps));
    std::memcpy(ctx->sigma, sigma, sizeof(ctx->sigma));
    std::memcpy(ctx->charge, charge, sizeof(ctx->charge));
    CG_HIP(ctx, hipSetDevice(ctx->device));
    int rc = build_tables(ctx);
    if (rc) return rc;
    ctx->have_ff = true;
    // residue codes carry the charged flag, and charges set the device
    // residue order: replicas must be set again
    if (ctx->have_rep) free_replicas(ctx);
    if (ctx->have_sys && (rc = build_order(ctx))) return rc;
    return CGPPD_OK;
}

int cgppd_set_box(cgppd_ctx *ctx, float box)
{
    if (!ctx) return CGPPD_E_ARG;
    if (!(box > 0.0f) || !std::isfinite(box)) return fail(ctx, CGPPD_E_ARG, "box must be a positive finite edge");
    ctx->box = box;
    ctx->have_box = true;
    ctx->mc_ready = false;
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_set_system(cgppd_ctx *ctx, int n_mol, const int *mol_off, const int *type, int n_noncrowder)
{
    if (!ctx) return CGPPD_E_ARG;
    if (n_mol < 1 || !mol_off || !type) return fail(ctx, CGPPD_E_ARG, "need n_mol >= 1 and non-null arrays");
    if (n_mol > CGPPD_MAX_MOLECULES)
        return fail(ctx, CGPPD_E_ARG, "at most %d molecules per system (got %d)", CGPPD_MAX_MOLECULES, n_mol);
    if (mol_off[0] != 0) return fail(ctx, CGPPD_E_ARG, "mol_off[0] must be 0");
    int max_n = 0;
    for (int m = 0; m < n_mol; ++m) {
        const int n = mol_off[m + 1] - mol_off[m];
        if (n < 1) return fail(ctx, CGPPD_E_ARG, "molecule %d has %d residues", m, n);
        max_n = std::max(max_n, n);
    }
    const int N = mol_off[n_mol];
    for (int i = 0; i < N; ++i)
        if (type[i] < 0 || type[i] >= CGPPD_AA_COUNT)
            return fail(ctx, CGPPD_E_ARG, "residue %d has type %d outside 0..19", i, type[i]);
    if (n_noncrowder < 0 || n_noncrowder > n_mol) return fail(ctx, CGPPD_E_ARG, "n_noncrowder out of range");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    free_replicas(ctx);
    ctx->M = n_mol;
    ctx->N = N;
    ctx->stride = (N + kJChunk - 1) / kJChunk * kJChunk;
    ctx->n_nc = n_noncrowder;
    ctx->max_n = max_n;
    ctx->mol_off.assign(mol_off, mol_off + n_mol + 1);
    ctx->type.assign(type, type + N);
    std::vector<int> mol_of(N);
    for (int m = 0; m < n_mol; ++m)
        for (int i = mol_off[m]; i < mol_off[m + 1]; ++i) mol_of[i] = m;
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->d_mol_off, n_mol + 1))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_mol_of, N))) return rc;
    CG_HIP(ctx, hipMemcpy(ctx->d_mol_off, mol_off, sizeof(int) * (n_mol + 1), hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_mol_of, mol_of.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    ctx->topo.assign(n_mol, cgppd_topo::Molecule());
    ctx->chain_index.assign(N, 0);
    ctx->topo_dirty = true;
    ctx->n_flex = 0;
    ctx->have_sys = true;
    if (ctx->have_ff && (rc = build_order(ctx))) return rc;
    return CGPPD_OK;
}

int cgppd_set_torsion_table(cgppd_ctx *ctx, const double *v, const double *delta)
{
    if (!ctx) return CGPPD_E_ARG;
    if (!v || !delta) return fail(ctx, CGPPD_E_ARG, "null torsion table");
    std::memcpy(ctx->tor_v, v, sizeof(ctx->tor_v));
    std::memcpy(ctx->tor_d, delta, sizeof(ctx->tor_d));
    ctx->topo_dirty = true;
    ctx->mc_ready = false;
    return CGPPD_OK;
}

static int read_segments(int n, int n_segments, const int *seg_off, const int *seg_res,
                         std::vector<std::vector<int>> &segs, std::string &err)
{
    if (n_segments < 0) {
        err = "negative segment count";
        return CGPPD_E_ARG;
    }
    if (n_segments > 0 && (!seg_off || !seg_res)) {
        err = "null segment arrays";
        return CGPPD_E_ARG;
    }
    for (int k = 0; k < n_segments; ++k) {
        if (seg_off[k + 1] < seg_off[k] || (k == 0 && seg_off[0] != 0)) {
            err = "segment offsets must start at 0 and not decrease";
            return CGPPD_E_ARG;
        }
        segs.emplace_back(seg_res + seg_off[k], seg_res + seg_off[k + 1]);
        for (int v : segs.back())
            if (v < 0 || v >= n) {
                err = "segment residue " + std::to_string(v) + " outside the molecule";
                return CGPPD_E_ARG;
            }
    }
    return CGPPD_OK;
}

int cgppd_set_topology(cgppd_ctx *ctx, int molecule, const int *chain_code, const int *chain_index, int all_flexible,
                       int n_segments, const int *seg_off, const int *seg_res)
{
    if (!ctx) return CGPPD_E_ARG;
    if (!ctx->have_sys) return fail(ctx, CGPPD_E_STATE, "system not set");
    if (molecule < 0 || molecule >= ctx->M) return fail(ctx, CGPPD_E_ARG, "molecule %d out of range", molecule);
    if (!chain_code || !chain_index) return fail(ctx, CGPPD_E_ARG, "null chain arrays");
    const int i0 = ctx->mol_off[molecule], n = ctx->mol_off[molecule + 1] - i0;
    std::vector<std::vector<int>> segs;
    std::string err;
    int rc = read_segments(n, n_segments, seg_off, seg_res, segs, err);
    if (rc) return fail(ctx, rc, "%s", err.c_str());
    int nchains = 0;
    for (int k = 0; k < n; ++k) {
        if (chain_index[k] < 0) return fail(ctx, CGPPD_E_ARG, "negative chain index");
        nchains = std::max(nchains, chain_index[k] + 1);
    }
    cgppd_topo::Molecule t;
    if (!cgppd_topo::build(n, chain_code, all_flexible != 0, segs, nchains, t, err))
        return fail(ctx, CGPPD_E_ARG, "%s", err.c_str());
    if (ctx->rep_crowd && molecule >= ctx->n_nc && t.flexible)
        return fail(ctx, CGPPD_E_ARG, "molecule %d is a crowder: repulsive crowding needs rigid crowders", molecule);
    if ((int)ctx->topo.size() != ctx->M) ctx->topo.assign(ctx->M, cgppd_topo::Molecule());
    ctx->topo[molecule] = std::move(t);
    std::copy(chain_index, chain_index + n, ctx->chain_index.begin() + i0);
    ctx->topo_dirty = true;
    ctx->mc_ready = false;
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_topology_list(int n_res, const int *chain_code, const int *chain_index, int all_flexible, int n_segments,
                        const int *seg_off, const int *seg_res, int kind, int *out, int *count)
{
    if (n_res < 1 || !chain_code || !chain_index || !count) return CGPPD_E_ARG;
    std::vector<std::vector<int>> segs;
    std::string err;
    if (read_segments(n_res, n_segments, seg_off, seg_res, segs, err)) return CGPPD_E_ARG;
    int nchains = 0;
    for (int k = 0; k < n_res; ++k) nchains = std::max(nchains, chain_index[k] + 1);
    cgppd_topo::Molecule t;
    if (!cgppd_topo::build(n_res, chain_code, all_flexible != 0, segs, nchains, t, err)) return CGPPD_E_ARG;
    std::vector<int> v;
    switch (kind) {
    case 0: for (const auto &a : t.bonds) v.insert(v.end(), a.begin(), a.end()); break;
    case 1: for (const auto &a : t.angles) v.insert(v.end(), a.begin(), a.end()); break;
    case 2: for (const auto &a : t.torsions) v.insert(v.end(), a.begin(), a.end()); break;
    case 3: v = t.local; break;
    case 4: v = t.crankshaft; break;
    case 5: v = t.flex; break;
    case 6: for (const auto &a : t.segment_bonds) v.insert(v.end(), a.begin(), a.end()); break;
    case 7: for (const auto &a : t.indirect) v.insert(v.end(), a.begin(), a.end()); break;
    case 8: v = t.domain_of; break;
    case 9: v = t.flex_nb; break;
    case 10: v = t.flex_nb_off; break;
    case 11: v = t.branch; break;
    case 12: v = t.branch_off; break;
    case 13: v = t.segbond_of; break;
    case 14: v = {t.vertex_count, t.num_chains}; break;
    default: return CGPPD_E_ARG;
    }
    *count = (int)v.size();
    if (out) std::copy(v.begin(), v.end(), out);
    return CGPPD_OK;
}

int cgppd_set_replicas(cgppd_ctx *ctx, int n_rep, const float *center, const float *rel, const double *temperature,
                       const float *translate_step, const float *rotate_step, const int *label, uint64_t seed)
{
    int rc = check_ready(ctx, false);
    if (rc) return rc;
    if (n_rep < 1 || !center || !rel) return fail(ctx, CGPPD_E_ARG, "need n_rep >= 1, centres and positions");
    if (n_rep > CGPPD_MAX_REPLICAS)
        return fail(ctx, CGPPD_E_ARG, "at most %d replicas per context (got %d)", CGPPD_MAX_REPLICAS, n_rep);
    CG_HIP(ctx, hipSetDevice(ctx->device));
    free_replicas(ctx);
    const int M = ctx->M, N = ctx->N, S = ctx->stride, R = n_rep;
    ctx->R = R;
    ctx->seed = seed;
    std::vector<float4> hpos((size_t)R * S), hrel((size_t)R * S), hcen((size_t)R * M);
    for (int r = 0; r < R; ++r) {
        for (int m = 0; m < M; ++m) {
            const float *c = center + ((size_t)r * M + m) * 3;
            hcen[(size_t)r * M + m] = make_float4(c[0], c[1], c[2], 0.0f);
            for (int i = ctx->mol_off[m]; i < ctx->mol_off[m + 1]; ++i) {
                const float *v = rel + ((size_t)r * N + ctx->perm[i]) * 3;
                const float w = code_bits(residue_word(ctx, i));
                hrel[(size_t)r * S + i] = make_float4(v[0], v[1], v[2], w);
                // src/Molecule.cpp:183-191: position = relative + centre
                hpos[(size_t)r * S + i] = make_float4(v[0] + c[0], v[1] + c[1], v[2] + c[2], w);
            }
        }
        for (int i = N; i < S; ++i) {
            hpos[(size_t)r * S + i] = make_float4(0.f, 0.f, 0.f, 0.f);
            hrel[(size_t)r * S + i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    std::vector<float> hrad((size_t)R * M);
    for (int r = 0; r < R; ++r)
        for (int m = 0; m < M; ++m)
            hrad[(size_t)r * M + m] = molecule_radius(&hrel[(size_t)r * S + ctx->mol_off[m]],
                                                      ctx->mol_off[m + 1] - ctx->mol_off[m]);
    std::vector<double> temp(R), rcos(R), rsin(R);
    std::vector<float> tstep(R);
    std::vector<int> lab(R), rank_of(R), at_rank(R);
    for (int r = 0; r < R; ++r) {
        temp[r] = temperature ? temperature[r] : 300.0;
        tstep[r] = translate_step ? translate_step[r] : 0.5f;
        const float rs = rotate_step ? rotate_step[r] : 0.2f;
        // src/Quaternion.cpp:14-15 with angle = rotate step (double)
        rcos[r] = std::cos((double)rs / 2.0);
        rsin[r] = std::sin((double)rs / 2.0);
        lab[r] = label ? label[r] : r;
        at_rank[r] = r;
    }
    std::stable_sort(at_rank.begin(), at_rank.end(), [&](int a, int b) { return temp[a] < temp[b]; });
    for (int k = 0; k < R; ++k) rank_of[at_rank[k]] = k;

    if ((rc = dev_alloc(ctx, &ctx->d_pos, (size_t)R * S))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_rel, (size_t)R * S))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_center, (size_t)R * M))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_radius, (size_t)R * M))) return rc;
    // FP64 rows, then the fixed-point rows (same size)
    if ((rc = dev_alloc(ctx, &ctx->d_rowbuf, (size_t)2 * R * M))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_emat, (size_t)R * M * M))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_energy, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_temp, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_rcos, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_rsin, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_tstep, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_done, R))) return rc;
    // 8 records of padding: the row kernel reads i records in groups of 8
    if ((rc = dev_alloc(ctx, &ctx->d_prop_pos, (size_t)R * ctx->max_n + 8))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_rel, (size_t)R * ctx->max_n))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_center, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_radius, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_moved, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_mol, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_nib, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_n, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_items, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_prop_hdr, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_counters, (size_t)3 * R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_steps, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_ex_round, 1))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_ex_stats, 2))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_pair_evals, 1))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_label, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_rank_of, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_rep_at_rank, R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_out, (size_t)2 * R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_out_int, (size_t)5 * R))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_out_scratch, (size_t)2 * R))) return rc;
    CG_HIP(ctx, hipHostMalloc((void **)&ctx->h_out, sizeof(double) * 2 * R, hipHostMallocDefault));

    CG_HIP(ctx, hipMemcpy(ctx->d_pos, hpos.data(), sizeof(float4) * hpos.size(), hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_rel, hrel.data(), sizeof(float4) * hrel.size(), hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_center, hcen.data(), sizeof(float4) * hcen.size(), hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_radius, hrad.data(), sizeof(float) * hrad.size(), hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_temp, temp.data(), sizeof(double) * R, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_rcos, rcos.data(), sizeof(double) * R, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_rsin, rsin.data(), sizeof(double) * R, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_tstep, tstep.data(), sizeof(float) * R, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_label, lab.data(), sizeof(int) * R, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_rank_of, rank_of.data(), sizeof(int) * R, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_rep_at_rank, at_rank.data(), sizeof(int) * R, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemset(ctx->d_rowbuf, 0, sizeof(double) * 2 * R * M));
    CG_HIP(ctx, hipMemset(ctx->d_emat, 0, sizeof(double) * R * M * M));
    CG_HIP(ctx, hipMemset(ctx->d_energy, 0, sizeof(double) * R));
    CG_HIP(ctx, hipMemset(ctx->d_done, 0, sizeof(unsigned) * R));
    CG_HIP(ctx, hipMemset(ctx->d_counters, 0, sizeof(long long) * 3 * R));
    CG_HIP(ctx, hipMemset(ctx->d_steps, 0, sizeof(long long) * R));
    CG_HIP(ctx, hipMemset(ctx->d_ex_round, 0, sizeof(long long)));
    CG_HIP(ctx, hipMemset(ctx->d_ex_stats, 0, sizeof(unsigned long long) * 2));
    CG_HIP(ctx, hipMemset(ctx->d_pair_evals, 0, sizeof(unsigned long long)));
    CG_HIP(ctx, hipMemset(ctx->d_out, 0, sizeof(double) * 2 * R));
    if ((rc = set_anchors(ctx, rel))) return rc;
    ctx->have_rep = true;
    ctx->mc_ready = false;
    return CGPPD_OK;
}

static int upload_molecules(cgppd_ctx *ctx, int replica, int m0, int m1, const float *xyz)
{
    const int i0 = ctx->mol_off[m0], n = ctx->mol_off[m1] - i0;
    std::vector<float4> hp(n), hr(n), hc(m1 - m0);
    std::vector<float> hrad(m1 - m0);
    for (int m = m0; m < m1; ++m) {
        const int a = ctx->mol_off[m] - i0, b = ctx->mol_off[m + 1] - i0;
        split_molecule(ctx, i0 + a, b - a, xyz + 3 * a, &hp[a], &hr[a], hc[m - m0], hrad[m - m0]);
    }
    const size_t rs = (size_t)replica * ctx->stride + i0, rm = (size_t)replica * ctx->M + m0;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipMemcpyAsync(ctx->d_pos + rs, hp.data(), sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream));
    CG_HIP(ctx, hipMemcpyAsync(ctx->d_rel + rs, hr.data(), sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream));
    CG_HIP(ctx, hipMemcpyAsync(ctx->d_center + rm, hc.data(), sizeof(float4) * (m1 - m0), hipMemcpyHostToDevice,
                               ctx->stream));
    CG_HIP(ctx, hipMemcpyAsync(ctx->d_radius + rm, hrad.data(), sizeof(float) * (m1 - m0), hipMemcpyHostToDevice,
                               ctx->stream));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->mc_ready = false;
    ctx->smat_dirty = true;
    return CGPPD_OK;
}

int cgppd_upload_positions(cgppd_ctx *ctx, int replica, const float *xyz)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (replica < 0 || replica >= ctx->R || !xyz) return fail(ctx, CGPPD_E_ARG, "bad replica or null coordinates");
    return upload_molecules(ctx, replica, 0, ctx->M, xyz);
}

int cgppd_update_molecule(cgppd_ctx *ctx, int replica, int molecule, const float *xyz)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (replica < 0 || replica >= ctx->R || molecule < 0 || molecule >= ctx->M || !xyz)
        return fail(ctx, CGPPD_E_ARG, "bad replica/molecule or null coordinates");
    return upload_molecules(ctx, replica, molecule, molecule + 1, xyz);
}

int cgppd_set_repulsive_crowding(cgppd_ctx *ctx, int enable)
{
    if (!ctx) return CGPPD_E_ARG;
    // (the topology each cgppd_set_topology call stores, not the device
    // lists, which are rebuilt lazily)
    if (enable && flexible_crowder(ctx) >= 0)
        return fail(ctx, CGPPD_E_ARG, "repulsive crowding needs rigid crowders (molecule %d is flexible)",
                    flexible_crowder(ctx));
    if (ctx->rep_crowd != (enable != 0)) {
        // the energy model changed: the MC matrix and totals must be rebuilt
        ctx->rep_crowd = enable != 0;
        ctx->mc_ready = false;
        ctx->smat_dirty = true;
    }
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_set_rigid_potential_only(cgppd_ctx *ctx, int enable)
{
    if (!ctx) return CGPPD_E_ARG;
    if (ctx->rigid_only != (enable != 0)) {
        ctx->rigid_only = enable != 0;
        ctx->mc_ready = false;
    }
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_set_image_shift(cgppd_ctx *ctx, int enable)
{
    if (!ctx) return CGPPD_E_ARG;
    ctx->fast_image = enable != 0;
    destroy_graph(ctx);
    return CGPPD_OK;
}

static int energy_common(cgppd_ctx *ctx, int replica, int n_active, bool intra)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (replica < 0 || replica >= ctx->R) return fail(ctx, CGPPD_E_ARG, "replica %d out of range", replica);
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipMemsetAsync(ctx->d_out + replica, 0, sizeof(double), ctx->stream));
    CG_HIP(ctx, hipMemsetAsync(ctx->d_out + ctx->R + replica, 0, sizeof(double), ctx->stream));
    if ((rc = launch_full(ctx, replica, 1, n_active, false, nullptr))) return rc;
    // flexible molecules' intra-molecular pairs, as EonDevice counts them
    // (the non-crowder variant counts inter-molecular pairs only,
    // src/CudaFunctions.cu:757-961)
    if (intra && (rc = launch_intra(ctx, replica, 1, false, nullptr, ctx->d_out, ctx->d_out + ctx->R))) return rc;
    return CGPPD_OK;
}

int cgppd_energy_internal(cgppd_ctx *ctx, int replica, double *out)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (replica < 0 || replica >= ctx->R || !out) return fail(ctx, CGPPD_E_ARG, "bad replica or null output");
    for (int k = 0; k < 5; ++k) out[k] = 0.0;
    if (ctx->n_flex == 0 || ctx->rigid_only) return CGPPD_OK;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = zero_dev(ctx, ctx->d_out_int, (size_t)5 * ctx->R))) return rc;
    if ((rc = launch_intra(ctx, replica, 1, false, nullptr, ctx->d_out_scratch, ctx->d_out_scratch + ctx->R)))
        return rc;
    double t[5];
    CG_HIP(ctx, hipMemcpyAsync(t, ctx->d_out_int + (size_t)5 * replica, sizeof(t), hipMemcpyDeviceToHost,
                               ctx->stream));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // src/Potential.cpp:203-210
    out[0] = t[0] * (double)ctx->rt_to_kcal;
    out[1] = t[1] * (double)ctx->dh_factor;
    out[2] = t[2] * 0.5 * 378.0;
    out[3] = t[3] * -10.0;
    out[4] = t[4];
    return CGPPD_OK;
}

int cgppd_energy(cgppd_ctx *ctx, int replica, double *lj, double *dh, double *total)
{
    int rc = energy_common(ctx, replica, ctx ? ctx->M : 0, true);
    if (rc) return rc;
    return energy_outputs(ctx, replica, lj, dh, total);
}

int cgppd_energy_nc(cgppd_ctx *ctx, int replica, double *lj, double *dh, double *total)
{
    int rc = energy_common(ctx, replica, ctx ? ctx->n_nc : 0, false);
    if (rc) return rc;
    return energy_outputs(ctx, replica, lj, dh, total);
}

int cgppd_energy_nc_all(cgppd_ctx *ctx, double *total)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!total) return fail(ctx, CGPPD_E_ARG, "null output");
    const int R = ctx->R;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = zero_dev(ctx, ctx->d_out, (size_t)2 * R))) return rc;
    if ((rc = launch_full(ctx, 0, R, ctx->n_nc, false, nullptr))) return rc;
    CG_HIP(ctx, hipMemcpyAsync(ctx->h_out, ctx->d_out, sizeof(double) * 2 * R, hipMemcpyDeviceToHost, ctx->stream));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // src/Potential.cpp:219
    for (int r = 0; r < R; ++r) total[r] = ctx->h_out[R + r] * (double)ctx->dh_factor + ctx->h_out[r] * (double)ctx->rt_to_kcal;
    return CGPPD_OK;
}

int cgppd_energy_async(cgppd_ctx *ctx, int replica) { return energy_common(ctx, replica, ctx ? ctx->M : 0, true); }

int cgppd_energy_wait(cgppd_ctx *ctx, int replica, double *lj, double *dh, double *total)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (replica < 0 || replica >= ctx->R) return fail(ctx, CGPPD_E_ARG, "replica %d out of range", replica);
    return energy_outputs(ctx, replica, lj, dh, total);
}

int cgppd_energy_all(cgppd_ctx *ctx, double *lj, double *dh, double *total, double *mat)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    const int R = ctx->R, M = ctx->M;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if (mat && !ctx->d_out_mat)
        if ((rc = dev_alloc(ctx, &ctx->d_out_mat, (size_t)R * M * M))) return rc;
    if ((rc = zero_dev(ctx, ctx->d_out, (size_t)2 * R))) return rc;
    if (mat && (rc = zero_dev(ctx, ctx->d_out_mat, (size_t)R * M * M))) return rc;
    if ((rc = launch_full(ctx, 0, R, M, false, mat ? ctx->d_out_mat : nullptr))) return rc;
    if ((rc = launch_intra(ctx, 0, R, false, mat ? ctx->d_out_mat : nullptr, ctx->d_out, ctx->d_out + R))) return rc;
    CG_HIP(ctx, hipMemcpyAsync(ctx->h_out, ctx->d_out, sizeof(double) * 2 * R, hipMemcpyDeviceToHost, ctx->stream));
    if (mat)
        CG_HIP(ctx, hipMemcpyAsync(mat, ctx->d_out_mat, sizeof(double) * R * M * M, hipMemcpyDeviceToHost, ctx->stream));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < R; ++r) {
        const double l = ctx->h_out[r], d = ctx->h_out[R + r];
        if (lj) lj[r] = l;
        if (dh) dh[r] = d;
        if (total) total[r] = d * (double)ctx->dh_factor + l * (double)ctx->rt_to_kcal;
    }
    if (mat)
        for (int r = 0; r < R; ++r) {
            double *m = mat + (size_t)r * M * M;
            for (int I = 0; I < M; ++I)
                for (int J = I + 1; J < M; ++J) m[(size_t)J * M + I] = m[(size_t)I * M + J];
        }
    return CGPPD_OK;
}

int cgppd_energy_all_noalloc(cgppd_ctx *ctx, double *lj, double *dh, double *total, double *mat)
{
    // nothing here may allocate: no lazy topology upload (check_ready would
    // do one), no context buffers created on first use, no host copies
    if (!ctx) return CGPPD_E_ARG;
    if (!ctx->have_ff || !ctx->have_box || !ctx->have_sys || !ctx->have_rep)
        return fail(ctx, CGPPD_E_STATE, "force field, box, system and replicas must be set");
    if (ctx->topo_dirty)
        return fail(ctx, CGPPD_E_STATE, "topology changed since the replicas were set: call cgppd_mc_prepare or "
                                        "cgppd_energy once first");
    const int R = ctx->R, M = ctx->M;
    if (R > CGPPD_NOALLOC_MAX_REPLICAS)
        return fail(ctx, CGPPD_E_ARG, "cgppd_energy_all_noalloc takes at most %d replicas (context has %d)",
                    CGPPD_NOALLOC_MAX_REPLICAS, R);
    double *acc = scratch_of(ctx->device);
    if (!acc) return fail(ctx, CGPPD_E_STATE, "no scratch address for device %d", ctx->device);
    CG_HIP(ctx, hipSetDevice(ctx->device));
    int rc;
    // the scratch keeps the last call's sums, and the matrix is summed into
    if ((rc = zero_dev(ctx, acc, (size_t)2 * R))) return rc;
    if (mat && (rc = zero_dev(ctx, mat, (size_t)R * M * M))) return rc;
    if ((rc = launch_full(ctx, 0, R, M, false, mat, acc, acc + R))) return rc;
    if ((rc = launch_intra(ctx, 0, R, false, mat, acc, acc + R))) return rc;
    const size_t n = std::max<size_t>((size_t)R, mat ? (size_t)R * M * M : 0);
    hipLaunchKernelGGL(noalloc_energy_finish_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, acc,
                       R, M, ctx->rt_to_kcal, ctx->dh_factor, lj, dh, total, mat);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

int cgppd_mc_prepare(cgppd_ctx *ctx)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = zero_dev(ctx, ctx->d_emat, (size_t)ctx->R * ctx->M * ctx->M))) return rc;
    if ((rc = launch_full(ctx, 0, ctx->R, ctx->M, true, nullptr))) return rc;
    if ((rc = launch_intra(ctx, 0, ctx->R, true, nullptr, nullptr, nullptr))) return rc;
    // the most i-lane items one proposal can have (ilane_items): systems
    // with flexible molecules size the pair kernel's grid by it
    {
        const int JC = (ctx->N + kJChunk - 1) / kJChunk;
        int most = 0;
        for (int m = 0; m < ctx->M; ++m) {
            const int n = ctx->mol_off[m + 1] - ctx->mol_off[m], nib = (n + kIBlock - 1) / kIBlock;
            int it = ctx->n_flex && ctx->N == n ? 0 : nib * JC;
            if (!ctx->rigid_only && ctx->n_flex && ctx->mol_flex_host[m])
                it += std::max((n + kJChunk - 1) / kJChunk, (intra_units(n) + kWavesPerWG - 1) / kWavesPerWG);
            most = std::max(most, it);
        }
        ctx->ilane_max_items = most;
    }
    PairParams p = base_params(ctx);
    hipLaunchKernelGGL(prepare_finish_kernel, dim3(ctx->R), dim3(kWG), 0, ctx->stream, p);
    CG_HIP(ctx, hipGetLastError());
    if ((rc = launch_finish(ctx, 1))) return rc;
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->mc_ready = true;
    ctx->smat_dirty = true;
    ctx->s1_valid = false;
    return CGPPD_OK;
}

static int mc_steps_launch(cgppd_ctx *ctx, int n_steps, int W);

static int ensure_events(cgppd_ctx *ctx, size_t n)
{
    while (ctx->events.size() < n) {
        hipEvent_t ev;
        CG_HIP(ctx, hipEventCreate(&ev));
        ctx->events.push_back(ev);
    }
    return CGPPD_OK;
}

int cgppd_mc_steps(cgppd_ctx *ctx, int n_steps)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!ctx->mc_ready) return fail(ctx, CGPPD_E_STATE, "call cgppd_mc_prepare before MC steps");
    if (n_steps < 0) return fail(ctx, CGPPD_E_ARG, "negative step count");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_stage1(ctx))) return rc;
    const int W = window_of(ctx);
    if ((rc = ensure_windows(ctx, W))) return rc;
    rc = mc_steps_launch(ctx, n_steps, W);
    if (rc == CGPPD_OK) ctx->h_step += n_steps;
    return rc;
}

// The launches of cgppd_mc_steps (the caller advances ctx->h_step).
static int mc_steps_launch(cgppd_ctx *ctx, int n_steps, int W)
{
    int rc;
    if (ctx->timing && n_steps > 0 && W > 1) {
        // the row kernel of each window launch pair
        const int L = windows_needed(ctx, ctx->h_step, n_steps, W);
        if ((rc = ensure_events(ctx, (size_t)2 * L))) return rc;
        if ((rc = launch_window_begin(ctx, n_steps))) return rc;
        for (int k = 0; k < L; ++k)
            if ((rc = launch_window_pair(ctx, ctx->events[2 * k], ctx->events[2 * k + 1]))) return rc;
        CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double sum = 0.0;
        for (int k = 0; k < L; ++k) {
            float ms = 0.f;
            CG_HIP(ctx, hipEventElapsedTime(&ms, ctx->events[2 * k], ctx->events[2 * k + 1]));
            sum += ms;
        }
        ctx->t_mean_ms = L ? sum / L : 0.0;
        ctx->t_launches = L;
        return CGPPD_OK;
    }
    if (ctx->timing && n_steps > 0 && resident_active(ctx)) {
        // the resident kernel runs the steps in ceil(n_steps / kResidentChunk)
        // launches: their span on the context stream, per MC step
        if ((rc = ensure_events(ctx, 2))) return rc;
        CG_HIP(ctx, hipEventRecord(ctx->events[0], ctx->stream));
        if ((rc = launch_resident_steps(ctx, n_steps))) return rc;
        CG_HIP(ctx, hipEventRecord(ctx->events[1], ctx->stream));
        CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        CG_HIP(ctx, hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
        ctx->t_mean_ms = (double)ms / n_steps;
        ctx->t_launches = (n_steps + kResidentChunk - 1) / kResidentChunk;
        return CGPPD_OK;
    }
    if (ctx->timing && n_steps > 0) {
        if ((rc = ensure_events(ctx, (size_t)2 * n_steps))) return rc;
        for (int k = 0; k < n_steps; ++k) {
            if ((rc = launch_pair(ctx, ctx->events[2 * k], ctx->events[2 * k + 1]))) return rc;
            if ((rc = launch_finish(ctx, 0))) return rc;
        }
        CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double sum = 0.0;
        for (int k = 0; k < n_steps; ++k) {
            float ms = 0.f;
            CG_HIP(ctx, hipEventElapsedTime(&ms, ctx->events[2 * k], ctx->events[2 * k + 1]));
            sum += ms;
        }
        ctx->t_mean_ms = sum / n_steps;
        ctx->t_launches = n_steps;
        return CGPPD_OK;
    }
    return launch_mc_steps(ctx, n_steps);
}

int cgppd_exchange(cgppd_ctx *ctx)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    ctx->s1_valid = false;
    if (!ctx->mc_ready) return fail(ctx, CGPPD_E_STATE, "call cgppd_mc_prepare before exchanges");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    return launch_exchange(ctx);
}

// Exchange intervals in speculative windows: the graph holds the run's
// start and graph_L window launch pairs; an interval whose replicas need
// more windows (counted on the host from the molecule draws) gets the rest
// as direct launches, and one that needs fewer leaves the graph's last
// pairs without work.  The exchange follows each interval.
static int run_windows_graph(cgppd_ctx *ctx, int n_intervals, int n, int W)
{
    int rc = ensure_windows(ctx, W);
    if (rc) return rc;
    if (!ctx->graph_exec || ctx->graph_steps != n || ctx->graph_win != W || ctx->graph_groups != 1) {
        destroy_graph(ctx);
        // the fewest windows of the first intervals of this run (at most 4)
        int L = 0;
        for (int k = 0; k < std::max(1, std::min(n_intervals, 4)); ++k) {
            const int l = windows_needed(ctx, ctx->h_step + (long long)k * n, n, W);
            L = k == 0 ? l : std::min(L, l);
        }
        L = std::max(1, L - 1);
        CG_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int lrc = launch_window_begin(ctx, n);
        for (int k = 0; k < L && !lrc; ++k) lrc = launch_window_pair(ctx);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (lrc) {
            if (g) (void)hipGraphDestroy(g);
            return lrc;
        }
        if (ce != hipSuccess) return fail(ctx, CGPPD_E_HIP, "graph capture: %s", hipGetErrorString(ce));
        ctx->graph = g;
        CG_HIP(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
        ctx->graph_steps = n;
        ctx->graph_groups = 1;
        ctx->graph_win = W;
        ctx->graph_L = L;
    }
    for (int k = 0; k < n_intervals; ++k) {
        // the graph's launches do not depend on the interval's window count:
        // the host replays the draws while the graph runs
        CG_HIP(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
        const int need = windows_needed(ctx, ctx->h_step, n, W);
        for (int e = ctx->graph_L; e < need; ++e)
            if ((rc = launch_window_pair(ctx))) return rc;
        if ((rc = launch_exchange(ctx))) return rc;
        ctx->h_step += n;
    }
    return CGPPD_OK;
}

int cgppd_run(cgppd_ctx *ctx, int n_intervals, int steps_per_interval, int use_graph)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!ctx->mc_ready) return fail(ctx, CGPPD_E_STATE, "call cgppd_mc_prepare before running");
    if (n_intervals < 0 || steps_per_interval < 0) return fail(ctx, CGPPD_E_ARG, "negative counts");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_stage1(ctx))) return rc;
    if (!use_graph) {
        for (int k = 0; k < n_intervals; ++k) {
            if ((rc = cgppd_mc_steps(ctx, steps_per_interval))) return rc;
            if ((rc = launch_exchange(ctx))) return rc;
            // exchanged temperatures change the stage-one test
            if ((rc = launch_stage1(ctx))) return rc;
        }
        return CGPPD_OK;
    }
    const int W = window_of(ctx);
    if (W > 1) return run_windows_graph(ctx, n_intervals, steps_per_interval, W);
    const int G = groups_of(ctx);
    if (!ctx->graph_exec || ctx->graph_steps != steps_per_interval || ctx->graph_groups != G || ctx->graph_win != 1) {
        destroy_graph(ctx);
        if ((rc = ensure_group_streams(ctx, G))) return rc;
        CG_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int lrc = launch_mc_steps(ctx, steps_per_interval);
        if (!lrc) lrc = launch_exchange(ctx);
        if (!lrc) lrc = launch_stage1(ctx);  // exchanged temperatures change the stage-one test
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (lrc) {
            if (g) (void)hipGraphDestroy(g);
            return lrc;
        }
        if (ce != hipSuccess) return fail(ctx, CGPPD_E_HIP, "graph capture: %s", hipGetErrorString(ce));
        ctx->graph = g;
        CG_HIP(ctx, hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0));
        ctx->graph_steps = steps_per_interval;
        ctx->graph_groups = G;
        ctx->graph_win = 1;
    }
    for (int k = 0; k < n_intervals; ++k) CG_HIP(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
    ctx->h_step += (long long)n_intervals * steps_per_interval;
    return CGPPD_OK;
}

int cgppd_synchronize(cgppd_ctx *ctx)
{
    if (!ctx) return CGPPD_E_ARG;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CGPPD_OK;
}

int cgppd_get_replica_stats(cgppd_ctx *ctx, double *energy, double *temperature, int *label, int64_t *counters,
                            int64_t *steps, int64_t *exchange)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    const int R = ctx->R;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (energy) CG_HIP(ctx, hipMemcpy(energy, ctx->d_energy, sizeof(double) * R, hipMemcpyDeviceToHost));
    if (temperature) CG_HIP(ctx, hipMemcpy(temperature, ctx->d_temp, sizeof(double) * R, hipMemcpyDeviceToHost));
    if (label) CG_HIP(ctx, hipMemcpy(label, ctx->d_label, sizeof(int) * R, hipMemcpyDeviceToHost));
    if (counters) CG_HIP(ctx, hipMemcpy(counters, ctx->d_counters, sizeof(long long) * 3 * R, hipMemcpyDeviceToHost));
    if (steps) CG_HIP(ctx, hipMemcpy(steps, ctx->d_steps, sizeof(long long) * R, hipMemcpyDeviceToHost));
    if (exchange) CG_HIP(ctx, hipMemcpy(exchange, ctx->d_ex_stats, sizeof(long long) * 2, hipMemcpyDeviceToHost));
    return CGPPD_OK;
}

int cgppd_get_pair_evaluations(cgppd_ctx *ctx, int64_t *pairs)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!pairs) return fail(ctx, CGPPD_E_ARG, "null output");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CG_HIP(ctx, hipMemcpy(pairs, ctx->d_pair_evals, sizeof(int64_t), hipMemcpyDeviceToHost));
    return CGPPD_OK;
}

int cgppd_get_coordinates(cgppd_ctx *ctx, int replica, float *pos, float *rel, float *center)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (replica < 0 || replica >= ctx->R) return fail(ctx, CGPPD_E_ARG, "replica %d out of range", replica);
    const int N = ctx->N, M = ctx->M;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<float4> h(std::max(N, M));
    if (pos) {
        CG_HIP(ctx, hipMemcpy(h.data(), ctx->d_pos + (size_t)replica * ctx->stride, sizeof(float4) * N,
                              hipMemcpyDeviceToHost));
        for (int i = 0; i < N; ++i) {
            const int u = ctx->perm[i];
            pos[3 * u] = h[i].x;
            pos[3 * u + 1] = h[i].y;
            pos[3 * u + 2] = h[i].z;
        }
    }
    if (rel) {
        CG_HIP(ctx, hipMemcpy(h.data(), ctx->d_rel + (size_t)replica * ctx->stride, sizeof(float4) * N,
                              hipMemcpyDeviceToHost));
        for (int i = 0; i < N; ++i) {
            const int u = ctx->perm[i];
            rel[3 * u] = h[i].x;
            rel[3 * u + 1] = h[i].y;
            rel[3 * u + 2] = h[i].z;
        }
    }
    if (center) {
        CG_HIP(ctx, hipMemcpy(h.data(), ctx->d_center + (size_t)replica * M, sizeof(float4) * M,
                              hipMemcpyDeviceToHost));
        for (int m = 0; m < M; ++m) {
            center[3 * m] = h[m].x;
            center[3 * m + 1] = h[m].y;
            center[3 * m + 2] = h[m].z;
        }
    }
    return CGPPD_OK;
}

int cgppd_get_energy_matrix(cgppd_ctx *ctx, int replica, double *mat)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (replica < 0 || replica >= ctx->R || !mat) return fail(ctx, CGPPD_E_ARG, "bad replica or null output");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CG_HIP(ctx, hipMemcpy(mat, ctx->d_emat + (size_t)replica * ctx->M * ctx->M,
                          sizeof(double) * ctx->M * ctx->M, hipMemcpyDeviceToHost));
    return CGPPD_OK;
}

int cgppd_set_launch(cgppd_ctx *ctx, int wg_per_cu)
{
    if (!ctx || wg_per_cu < 0 || wg_per_cu > 64) return CGPPD_E_ARG;
    ctx->wg_per_cu = wg_per_cu;
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_set_replica_groups(cgppd_ctx *ctx, int groups)
{
    if (!ctx || groups < 0 || groups > kMaxGroups) return CGPPD_E_ARG;
    ctx->replica_groups = groups;
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_get_replica_groups(cgppd_ctx *ctx, int *groups)
{
    if (!ctx || !groups) return CGPPD_E_ARG;
    *groups = groups_of(ctx);
    return CGPPD_OK;
}

int cgppd_set_pair_layout(cgppd_ctx *ctx, int layout)
{
    if (!ctx || layout < CGPPD_LAYOUT_AUTO || layout > CGPPD_LAYOUT_RESIDENT) return CGPPD_E_ARG;
    ctx->pair_layout = layout;
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_set_window(cgppd_ctx *ctx, int slots)
{
    if (!ctx) return CGPPD_E_ARG;
    if (slots != 0 && slots != 1 && slots != 2 && slots != 4 && slots != 8)
        return fail(ctx, CGPPD_E_ARG, "window slots must be 0 (automatic), 1 (off), 2, 4 or 8 (got %d)", slots);
    if (slots != ctx->window) destroy_graph(ctx);
    ctx->window = slots;
    return CGPPD_OK;
}

int cgppd_get_window(cgppd_ctx *ctx, int *slots)
{
    if (!ctx || !slots) return CGPPD_E_ARG;
    *slots = ctx->have_sys ? window_of(ctx) : 1;
    return CGPPD_OK;
}

int cgppd_get_pair_launch(cgppd_ctx *ctx, int *layout, int *wg_per_cu)
{
    if (!ctx) return CGPPD_E_ARG;
    const int l = pair_layout_of(ctx);
    if (layout) *layout = ctx->have_sys && resident_active(ctx) ? CGPPD_LAYOUT_RESIDENT : l;
    if (wg_per_cu) *wg_per_cu = pair_wg_per_cu(ctx, l);
    return CGPPD_OK;
}

int cgppd_autotune_table(cgppd_ctx *ctx, double *ms)
{
    if (!ctx || !ms) return CGPPD_E_ARG;
    for (int l = 0; l < 3; ++l)
        for (int w = 0; w < 7; ++w) ms[l * 7 + w] = ctx->tune_ms[l][w];
    return CGPPD_OK;
}

int cgppd_autotune(cgppd_ctx *ctx, int launches, int *layout, int *wg_per_cu, double *best_ms)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!ctx->mc_ready) return fail(ctx, CGPPD_E_STATE, "call cgppd_mc_prepare before tuning");
    if (launches < 1 || launches > 100000) return fail(ctx, CGPPD_E_ARG, "launches must be in [1, 100000]");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_stage1(ctx))) return rc;
    if ((rc = ensure_events(ctx, (size_t)2 * launches))) return rc;
    // the timed launches add into the pending rows, which already hold the
    // proposals' starting values (a flexible molecule's internal energy):
    // keep a copy to put back
    const size_t row_n = (size_t)2 * ctx->R * ctx->M;
    double *row_copy = nullptr;
    CG_HIP(ctx, hipMalloc((void **)&row_copy, row_n * sizeof(double)));
    if (hipMemcpyAsync(row_copy, ctx->d_rowbuf, row_n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream) !=
        hipSuccess) {
        (void)hipFree(row_copy);
        return fail(ctx, CGPPD_E_HIP, "autotune: saving the rows failed");
    }
    // the caller's settings are lifted while candidates are timed
    const int keep_layout = ctx->pair_layout, keep_wg = ctx->wg_per_cu;
    const int layouts[3] = {CGPPD_LAYOUT_JLANES, CGPPD_LAYOUT_JLANES4, CGPPD_LAYOUT_ILANES};
    const int n_layouts = ctx->n_flex ? 1 : 3;
    int best_l = 0, best_w = 0;
    double best = 1e300;
    for (auto &row : ctx->tune_ms)
        for (double &v : row) v = 0.0;
    bool windowed = false;
    const int keep_window = ctx->window;
    int best_win = 0;
    for (int li = 0; li < n_layouts && rc == CGPPD_OK; ++li) {
        ctx->pair_layout = ctx->n_flex ? CGPPD_LAYOUT_AUTO : layouts[li];
        // window candidates: with the automatic choice, 1, 4 and 8 slots
        // where the layout runs windows, else the caller's
        int cands[3] = {0, 0, 0}, n_cands = 0;
        ctx->window = keep_window == 0 ? 4 : keep_window;
        if (keep_window == 0 && window_of(ctx) > 1) {
            cands[0] = 1;
            cands[1] = 4;
            cands[2] = 8;
            n_cands = 3;
        } else {
            cands[0] = keep_window;
            n_cands = 1;
        }
        for (int ci = 0; ci < n_cands && rc == CGPPD_OK; ++ci) {
        ctx->window = cands[ci];
        // a layout that runs speculative windows is timed on the window
        // laid out from the current state; the time per MC step is the
        // launch time plus the decision kernel's (an estimate, from traces
        // of a 40-chain system: 6 us for one step, 15 us for a window),
        // divided by the mean window length of the next 1000 steps
        const int Wc = window_of(ctx);
        double steps_per_launch = 1.0;
        const double finish_ms = Wc > 1 ? 0.015 : 0.006;
        if (Wc > 1) {
            if ((rc = ensure_windows(ctx, Wc))) break;
            if ((rc = launch_window_begin(ctx, 1 << 20))) break;
            double windows = 0.0;
            for (int r = 0; r < ctx->R; ++r) windows += (double)windows_of_replica(ctx, r, ctx->h_step, 1000, Wc);
            steps_per_launch = 1000.0 * ctx->R / windows;
            windowed = true;
        }
        for (int w = 2; w <= 8 && rc == CGPPD_OK; ++w) {
            ctx->wg_per_cu = w;
            // one untimed launch, then the timed ones
            if ((rc = launch_pair(ctx))) break;
            for (int k = 0; k < launches; ++k)
                if ((rc = launch_pair(ctx, ctx->events[2 * k], ctx->events[2 * k + 1]))) break;
            if (rc) break;
            if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
                rc = fail(ctx, CGPPD_E_HIP, "autotune: stream synchronisation failed");
                break;
            }
            double sum = 0.0;
            for (int k = 0; k < launches; ++k) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, ctx->events[2 * k], ctx->events[2 * k + 1]) != hipSuccess) {
                    rc = fail(ctx, CGPPD_E_HIP, "autotune: event timing failed");
                    break;
                }
                sum += ms;
            }
            const double mean = (sum / launches + finish_ms) / steps_per_launch;
            double &slot = ctx->tune_ms[ctx->n_flex ? 2 : li][w - 2];
            if (slot == 0.0 || mean < slot) {
                slot = mean;
                ctx->tune_win[ctx->n_flex ? 2 : li][w - 2] = Wc;
            }
            if (rc == CGPPD_OK && mean < best) {
                best = mean;
                best_l = ctx->n_flex ? CGPPD_LAYOUT_ILANES : layouts[li];
                best_w = w;
                best_win = Wc;
            }
        }
        }
    }
    ctx->window = keep_window;
    // ties (within 1 % of the best, same layout) go to more workgroups per
    // CU: with the finish kernel between launches, the fuller grid measured
    // faster in replayed intervals (1.074 M against 1.058 M MC steps/s on a
    // 40-chain system at 7 and 4 per CU, which timed the same here)
    if (rc == CGPPD_OK && best_l) {
        const int li = best_l == CGPPD_LAYOUT_ILANES ? 2 : best_l == CGPPD_LAYOUT_JLANES4 ? 1 : 0;
        for (int w = 8; w > best_w; --w) {
            const double v = ctx->tune_ms[li][w - 2];
            if (v > 0.0 && v <= best * 1.01) {
                best_w = w;
                best = v;
                best_win = ctx->tune_win[li][w - 2];
                break;
            }
        }
    }
    ctx->pair_layout = keep_layout;
    ctx->wg_per_cu = keep_wg;
    // the timed window launches added into the window rows, which are zero
    // between windows
    if (windowed && rc == CGPPD_OK) {
        const size_t rows = (size_t)ctx->R * ctx->win_alloc * (ctx->M + ctx->win_alloc);
        if ((rc = zero_dev(ctx, ctx->d_wrowbuf, rows)) == CGPPD_OK)
            rc = zero_dev(ctx, reinterpret_cast<double *>(ctx->d_wrowfix), rows);
    }
    const hipError_t back =
        hipMemcpyAsync(ctx->d_rowbuf, row_copy, row_n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
    const hipError_t done = hipStreamSynchronize(ctx->stream);
    (void)hipFree(row_copy);
    if (rc) return rc;
    if (back != hipSuccess || done != hipSuccess) return fail(ctx, CGPPD_E_HIP, "autotune: restoring the rows failed");
    ctx->tuned_layout = best_l;
    ctx->tuned_wg = best_w;
    ctx->tuned_window = best_win;
    destroy_graph(ctx);
    if (layout) *layout = best_l;
    if (wg_per_cu) *wg_per_cu = best_w;
    if (best_ms) *best_ms = best;
    return CGPPD_OK;
}

int cgppd_set_timing(cgppd_ctx *ctx, int enable)
{
    if (!ctx) return CGPPD_E_ARG;
    ctx->timing = enable != 0;
    return CGPPD_OK;
}

int cgppd_last_kernel_ms(cgppd_ctx *ctx, double *mean_ms, int64_t *launches)
{
    if (!ctx) return CGPPD_E_ARG;
    if (mean_ms) *mean_ms = ctx->t_mean_ms;
    if (launches) *launches = ctx->t_launches;
    return CGPPD_OK;
}

int cgppd_get_sizes(const cgppd_ctx *ctx, int *n_mol, int *n_res, int *n_rep)
{
    if (!ctx) return CGPPD_E_ARG;
    if (n_mol) *n_mol = ctx->M;
    if (n_res) *n_res = ctx->N;
    if (n_rep) *n_rep = ctx->R;
    return CGPPD_OK;
}

// ---------------------------------------------------------------------------
// Surrogate
// ---------------------------------------------------------------------------
static int ensure_frames(cgppd_ctx *ctx)
{
    if (ctx->d_frames) return CGPPD_OK;
    return dev_alloc(ctx, &ctx->d_frames, (size_t)ctx->R * ctx->M * 9);
}

static int ensure_surrogate(cgppd_ctx *ctx)
{
    if (ctx->have_sur) return CGPPD_OK;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    int rc;
    if ((rc = dev_alloc(ctx, &ctx->d_sp, mlp::kParams)) || (rc = dev_alloc(ctx, &ctx->d_sm, mlp::kParams)) ||
        (rc = dev_alloc(ctx, &ctx->d_sv, mlp::kParams)) || (rc = dev_alloc(ctx, &ctx->d_sg, mlp::kParams)) ||
        (rc = dev_alloc(ctx, &ctx->d_swb, mlp::kBf16Params)) || (rc = dev_alloc(ctx, &ctx->d_sloss, 1)))
        return rc;
    CG_HIP(ctx, hipMemset(ctx->d_sp, 0, sizeof(float) * mlp::kParams));
    CG_HIP(ctx, hipMemset(ctx->d_sm, 0, sizeof(float) * mlp::kParams));
    CG_HIP(ctx, hipMemset(ctx->d_sv, 0, sizeof(float) * mlp::kParams));
    CG_HIP(ctx, hipMemset(ctx->d_sg, 0, sizeof(float) * mlp::kParams));
    CG_HIP(ctx, hipMemset(ctx->d_swb, 0, sizeof(unsigned short) * mlp::kBf16Params));
    ctx->have_sur = true;
    ctx->adam_t = 0;
    return CGPPD_OK;
}

static int pack_weights(cgppd_ctx *ctx)
{
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((mlp::kBf16Params + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_sp,
                       ctx->d_swb);
    CG_HIP(ctx, hipGetLastError());
    return CGPPD_OK;
}

// scratch rows for host-provided batches (features as BF16 bits)
static int stage_rows(cgppd_ctx *ctx, const float *x, const float *t, int n)
{
    if ((size_t)n > ctx->t_cap) {
        int rc;
        const size_t cap = std::max<size_t>((size_t)n, 4096);
        if ((rc = dev_alloc(ctx, &ctx->d_tx, cap * mlp::kFeat)) || (rc = dev_alloc(ctx, &ctx->d_tt, cap)) ||
            (rc = dev_alloc(ctx, &ctx->d_ty, cap)))
            return rc;
        ctx->t_cap = cap;
    }
    std::vector<unsigned short> hx((size_t)n * mlp::kFeat);
    for (size_t i = 0; i < hx.size(); ++i) hx[i] = mlp::f2bf(x[i]);
    CG_HIP(ctx, hipMemcpy(ctx->d_tx, hx.data(), sizeof(unsigned short) * hx.size(), hipMemcpyHostToDevice));
    if (t) CG_HIP(ctx, hipMemcpy(ctx->d_tt, t, sizeof(float) * n, hipMemcpyHostToDevice));
    return CGPPD_OK;
}

int cgppd_surrogate_set_params(cgppd_ctx *ctx, const float *params)
{
    if (!ctx || !params) return CGPPD_E_ARG;
    int rc = ensure_surrogate(ctx);
    if (rc) return rc;
    CG_HIP(ctx, hipMemcpy(ctx->d_sp, params, sizeof(float) * mlp::kParams, hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemset(ctx->d_sm, 0, sizeof(float) * mlp::kParams));
    CG_HIP(ctx, hipMemset(ctx->d_sv, 0, sizeof(float) * mlp::kParams));
    CG_HIP(ctx, hipMemset(ctx->d_sg, 0, sizeof(float) * mlp::kParams));
    ctx->adam_t = 0;
    if ((rc = pack_weights(ctx))) return rc;
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->smat_dirty = true;
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_surrogate_init(cgppd_ctx *ctx, uint64_t seed)
{
    if (!ctx) return CGPPD_E_ARG;
    // He-normal weights (Box-Muller over Philox uniforms), zero biases
    std::vector<float> p(mlp::kParams, 0.0f);
    auto normal = [&](uint32_t i) {
        uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, w[4];
        uint32_t c0 = i, c1 = 0, c2 = 0, c3 = kTagSG | 0xFFFFu;
        for (int round = 0; round < 10; ++round) {
            const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
            const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ key[0], n2 = (uint32_t)(p0 >> 32) ^ c3 ^ key[1];
            c0 = n0;
            c1 = (uint32_t)p1;
            c2 = n2;
            c3 = (uint32_t)p0;
            key[0] += 0x9E3779B9u;
            key[1] += 0xBB67AE85u;
        }
        w[0] = c0;
        w[1] = c1;
        const double u1 = ((double)w[0] + 1.0) / 4294967297.0, u2 = (double)w[1] / 4294967296.0;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < mlp::kFeat * mlp::kHidden; ++i)
        p[mlp::kOW1 + i] = (float)(normal(i) * std::sqrt(2.0 / mlp::kFeat));
    for (int i = 0; i < mlp::kHidden * mlp::kHidden; ++i)
        p[mlp::kOW2 + i] = (float)(normal(mlp::kOW2 + i) * std::sqrt(2.0 / mlp::kHidden));
    for (int i = 0; i < mlp::kHidden; ++i) p[mlp::kOw3 + i] = (float)(normal(mlp::kOw3 + i) * std::sqrt(1.0 / mlp::kHidden));
    return cgppd_surrogate_set_params(ctx, p.data());
}

int cgppd_surrogate_get_params(cgppd_ctx *ctx, float *params)
{
    if (!ctx || !params) return CGPPD_E_ARG;
    if (!ctx->have_sur) return fail(ctx, CGPPD_E_STATE, "surrogate not initialised");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CG_HIP(ctx, hipMemcpy(params, ctx->d_sp, sizeof(float) * mlp::kParams, hipMemcpyDeviceToHost));
    return CGPPD_OK;
}

int cgppd_surrogate_param_count(int *count)
{
    if (!count) return CGPPD_E_ARG;
    *count = mlp::kParams;
    return CGPPD_OK;
}

int cgppd_surrogate_predict(cgppd_ctx *ctx, const float *x, int n, float *y)
{
    if (!ctx || !x || !y || n < 0) return CGPPD_E_ARG;
    if (!ctx->have_sur) return fail(ctx, CGPPD_E_STATE, "surrogate not initialised");
    if (n == 0) return CGPPD_OK;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    int rc = stage_rows(ctx, x, nullptr, n);
    if (rc) return rc;
    const int g = (n + mlp::kRows - 1) / mlp::kRows;
    hipLaunchKernelGGL(mlp_predict_kernel, dim3(g), dim3(mlp::kThreads), 0, ctx->stream, ctx->d_tx, n, ctx->d_sp,
                       ctx->d_swb, ctx->d_ty);
    CG_HIP(ctx, hipGetLastError());
    CG_HIP(ctx, hipMemcpyAsync(y, ctx->d_ty, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CGPPD_OK;
}

int cgppd_surrogate_gradient(cgppd_ctx *ctx, const float *x, const float *t, int n, float *grad, double *loss)
{
    if (!ctx || !x || !t || !grad || n < 1) return CGPPD_E_ARG;
    if (!ctx->have_sur) return fail(ctx, CGPPD_E_STATE, "surrogate not initialised");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    int rc = stage_rows(ctx, x, t, n);
    if (rc) return rc;
    CG_HIP(ctx, hipMemsetAsync(ctx->d_sg, 0, sizeof(float) * mlp::kParams, ctx->stream));
    CG_HIP(ctx, hipMemsetAsync(ctx->d_sloss, 0, sizeof(double), ctx->stream));
    const int g = (n + mlp::kRows - 1) / mlp::kRows;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(g), dim3(mlp::kThreads), 0, ctx->stream, ctx->d_tx, ctx->d_tt, n, 1,
                       (uint64_t)0, (uint64_t)0, (float)n, ctx->d_sp, ctx->d_swb, ctx->d_sg, ctx->d_sloss);
    CG_HIP(ctx, hipGetLastError());
    CG_HIP(ctx, hipMemcpyAsync(grad, ctx->d_sg, sizeof(float) * mlp::kParams, hipMemcpyDeviceToHost, ctx->stream));
    double l = 0.0;
    CG_HIP(ctx, hipMemcpyAsync(&l, ctx->d_sloss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CG_HIP(ctx, hipMemsetAsync(ctx->d_sg, 0, sizeof(float) * mlp::kParams, ctx->stream));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l;
    return CGPPD_OK;
}

static int reserve_samples(cgppd_ctx *ctx, size_t need)
{
    if (need <= ctx->s_cap) return CGPPD_OK;
    const size_t cap = std::max(need, std::max<size_t>(ctx->s_cap * 2, (size_t)1 << 16));
    unsigned short *nx = nullptr;
    float *nt = nullptr;
    CG_HIP(ctx, hipMalloc((void **)&nx, sizeof(unsigned short) * cap * mlp::kFeat));
    CG_HIP(ctx, hipMalloc((void **)&nt, sizeof(float) * cap));
    if (ctx->s_n) {
        CG_HIP(ctx, hipMemcpyAsync(nx, ctx->d_sx, sizeof(unsigned short) * ctx->s_n * mlp::kFeat,
                                   hipMemcpyDeviceToDevice, ctx->stream));
        CG_HIP(ctx, hipMemcpyAsync(nt, ctx->d_st, sizeof(float) * ctx->s_n, hipMemcpyDeviceToDevice, ctx->stream));
        CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    dev_free(ctx->d_sx);
    dev_free(ctx->d_st);
    ctx->d_sx = nx;
    ctx->d_st = nt;
    ctx->s_cap = cap;
    return CGPPD_OK;
}

int cgppd_surrogate_add_samples(cgppd_ctx *ctx, const float *x, const float *t, int n)
{
    if (!ctx || !x || !t || n < 0) return CGPPD_E_ARG;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    int rc = reserve_samples(ctx, ctx->s_n + (size_t)n);
    if (rc) return rc;
    std::vector<unsigned short> hx((size_t)n * mlp::kFeat);
    for (size_t i = 0; i < hx.size(); ++i) hx[i] = mlp::f2bf(x[i]);
    CG_HIP(ctx, hipMemcpy(ctx->d_sx + ctx->s_n * mlp::kFeat, hx.data(), sizeof(unsigned short) * hx.size(),
                          hipMemcpyHostToDevice));
    CG_HIP(ctx, hipMemcpy(ctx->d_st + ctx->s_n, t, sizeof(float) * n, hipMemcpyHostToDevice));
    ctx->s_n += n;
    return CGPPD_OK;
}

int cgppd_surrogate_sample_count(cgppd_ctx *ctx, int64_t *count)
{
    if (!ctx || !count) return CGPPD_E_ARG;
    *count = (int64_t)ctx->s_n;
    return CGPPD_OK;
}

int cgppd_surrogate_clear_samples(cgppd_ctx *ctx)
{
    if (!ctx) return CGPPD_E_ARG;
    ctx->s_n = 0;
    return CGPPD_OK;
}

int cgppd_surrogate_train(cgppd_ctx *ctx, int steps, int batch, float lr, uint64_t seed, double *mean_loss)
{
    if (!ctx || steps < 0 || batch < mlp::kRows || batch % mlp::kRows) return fail(ctx, CGPPD_E_ARG, "batch must be a positive multiple of 32");
    if (!ctx->have_sur) return fail(ctx, CGPPD_E_STATE, "surrogate not initialised");
    if (ctx->s_n == 0) return fail(ctx, CGPPD_E_STATE, "no training samples");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipMemsetAsync(ctx->d_sloss, 0, sizeof(double), ctx->stream));
    const int g = batch / mlp::kRows;
    for (int k = 0; k < steps; ++k) {
        const uint64_t t = (uint64_t)++ctx->adam_t;
        hipLaunchKernelGGL(mlp_grad_kernel, dim3(g), dim3(mlp::kThreads), 0, ctx->stream, ctx->d_sx, ctx->d_st,
                           (int)ctx->s_n, 0, seed, t, (float)batch, ctx->d_sp, ctx->d_swb, ctx->d_sg, ctx->d_sloss);
        CG_HIP(ctx, hipGetLastError());
        const float bc1 = (float)(1.0 - std::pow(0.9, (double)t)), bc2 = (float)(1.0 - std::pow(0.999, (double)t));
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((mlp::kParams + 255) / 256), dim3(256), 0, ctx->stream, ctx->d_sp,
                           ctx->d_sm, ctx->d_sv, ctx->d_sg, lr, bc1, bc2);
        CG_HIP(ctx, hipGetLastError());
        int rc = pack_weights(ctx);
        if (rc) return rc;
    }
    double l = 0.0;
    CG_HIP(ctx, hipMemcpyAsync(&l, ctx->d_sloss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (mean_loss) *mean_loss = steps ? l / steps : 0.0;
    ctx->smat_dirty = true;
    destroy_graph(ctx);
    return CGPPD_OK;
}

int cgppd_surrogate_collect(cgppd_ctx *ctx)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!ctx->mc_ready) return fail(ctx, CGPPD_E_STATE, "call cgppd_mc_prepare before collecting samples");
    const int M = ctx->M, R = ctx->R;
    const size_t P = (size_t)M * (M - 1) / 2, n = P * R;
    if (n == 0) return CGPPD_OK;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_frames(ctx))) return rc;
    if ((rc = launch_frames(ctx))) return rc;
    // pairs within the surrogate's range, in (replica, I, J) order
    int *d_flag = nullptr;
    CG_HIP(ctx, hipMalloc((void **)&d_flag, sizeof(int) * n));
    PairParams p = base_params(ctx);
    hipLaunchKernelGGL(pair_near_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, d_flag);
    std::vector<int> slot(n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(slot.data(), d_flag, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(d_flag);
        return fail(ctx, CGPPD_E_HIP, "surrogate collect: %s", hipGetErrorString(e));
    }
    size_t kept = 0;
    for (size_t k = 0; k < n; ++k) slot[k] = slot[k] ? (int)(ctx->s_n + kept++) : -1;
    if ((rc = reserve_samples(ctx, ctx->s_n + kept))) {
        (void)hipFree(d_flag);
        return rc;
    }
    e = hipMemcpy(d_flag, slot.data(), sizeof(int) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess && kept > 0) {
        hipLaunchKernelGGL(pair_features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p,
                           ctx->d_sx, ctx->d_st, (float *)nullptr, (const int *)d_flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_flag);
    if (e != hipSuccess) return fail(ctx, CGPPD_E_HIP, "surrogate collect: %s", hipGetErrorString(e));
    ctx->s_n += kept;
    return CGPPD_OK;
}

int cgppd_surrogate_features(cgppd_ctx *ctx, int replica, float *x)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (replica < 0 || replica >= ctx->R || !x) return fail(ctx, CGPPD_E_ARG, "bad replica or null output");
    const int M = ctx->M;
    const size_t P = (size_t)M * (M - 1) / 2;
    if (P == 0) return CGPPD_OK;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    if ((rc = ensure_frames(ctx))) return rc;
    if ((rc = launch_frames(ctx))) return rc;
    float *d = nullptr;
    CG_HIP(ctx, hipMalloc((void **)&d, sizeof(float) * P * mlp::kFeat));
    PairParams p = base_params(ctx);
    p.rep_begin = replica;
    p.rep_count = 1;
    hipLaunchKernelGGL(pair_features_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, ctx->stream, p,
                       (unsigned short *)nullptr, (float *)nullptr, d, (const int *)nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(x, d, sizeof(float) * P * mlp::kFeat, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(ctx, CGPPD_E_HIP, "surrogate features: %s", hipGetErrorString(e));
    return CGPPD_OK;
}

int cgppd_set_delayed_acceptance(cgppd_ctx *ctx, int enable)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    CG_HIP(ctx, hipSetDevice(ctx->device));
    destroy_graph(ctx);
    if (!enable) {
        ctx->da = false;
        return CGPPD_OK;
    }
    if (!ctx->have_sur) return fail(ctx, CGPPD_E_STATE, "initialise or train the surrogate first");
    const size_t R = ctx->R, M = ctx->M;
    if (!ctx->d_smat) {
        if ((rc = dev_alloc(ctx, &ctx->d_smat, R * M * M)) || (rc = dev_alloc(ctx, &ctx->d_srow, R * M)) ||
            (rc = dev_alloc(ctx, &ctx->d_prop_frame, R * 9)) || (rc = dev_alloc(ctx, &ctx->d_s1, R)) ||
            (rc = dev_alloc(ctx, &ctx->d_s_dS, R)) || (rc = dev_alloc(ctx, &ctx->d_s1_rej, R)))
            return rc;
        CG_HIP(ctx, hipMemset(ctx->d_s1_rej, 0, sizeof(long long) * R));
        CG_HIP(ctx, hipMemset(ctx->d_s1, 0, sizeof(int) * R));
    }
    if ((rc = ensure_frames(ctx))) return rc;
    ctx->da = true;
    ctx->smat_dirty = true;
    ctx->s1_valid = false;
    return CGPPD_OK;
}

int cgppd_get_stage1_rejections(cgppd_ctx *ctx, int64_t *rejections)
{
    int rc = check_ready(ctx, true);
    if (rc) return rc;
    if (!rejections) return fail(ctx, CGPPD_E_ARG, "null output");
    CG_HIP(ctx, hipSetDevice(ctx->device));
    CG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!ctx->d_s1_rej) {
        for (int r = 0; r < ctx->R; ++r) rejections[r] = 0;
        return CGPPD_OK;
    }
    CG_HIP(ctx, hipMemcpy(rejections, ctx->d_s1_rej, sizeof(long long) * ctx->R, hipMemcpyDeviceToHost));
    return CGPPD_OK;
}

}  // extern "C""
