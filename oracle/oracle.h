/*
 * oracle.h -- CPU restatement of cgppd's REMC energy hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cgppd_amd/) links,
 * loads or calls this library; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker and as the timed CPU
 * baseline.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * each function is checked against) is listed beside it; functions with no
 * independent check say "parity unpinned".
 *
 * Layout conventions shared by all entry points:
 *   n_mol            molecules in the system
 *   mol_off[n_mol+1] first residue of each molecule (mol_off[n_mol] = N)
 *   type[N]          amino-acid index 0..19 (reference order A R N D C E Q G L K M F P S T W Y V H I)
 *   charge[20], sigma[20]  per-type charge and van der Waals radius
 *   eps[400]         raw contact energies e_ij [RT] (AminoAcids::LJpotentials)
 *   pos[N*3]         xyz of every residue, row-major
 *   box              periodic box edge (0 = no periodic boundary)
 */
#ifndef CGPPD_ORACLE_H
#define CGPPD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Constants as the reference evaluates them (src/constants.h:8-26), float. */
float orc_rt_to_kcal(void);      /* T_room*Rgas/kcal                       */
float orc_dh_factor(void);       /* e^2 N_A / (4 pi eps0 kcal Ang D_water)  */
float orc_r0_factor(void);       /* 2^(1/6), float pow                     */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Pinned by the
 * Random123 known-answer vectors in tests/test_oracle.py. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Reference-semantics inter-molecular energy: src/Replica.cpp:421-500 with
 * src/Residue.cpp:62-76 (FP32 minimum image + sqrtf), src/Potential.cpp:5-42
 * (double LJ, float expf DH), double accumulation.  Raw LJ [RT] and raw DH
 * are returned; total = dh*orc_dh_factor() + lj*orc_rt_to_kcal()
 * (src/Potential.cpp:214-221).  Pinned by the ten reference conformations
 * (tests/testReferenceConformations.cpp:83-94) and bit-exact against
 * oracle/_ref built from the reference sources. */
void orc_energy_ref(int n_mol, const int *mol_off, const int *type,
                    const float *charge, const float *sigma, const float *eps,
                    const float *pos, float box, double *lj, double *dh);

/* Same, restricted to molecules a and b: src/Replica.cpp:502-528. */
void orc_energy_pair_ref(const int *mol_off, int a, int b, const int *type,
                         const float *charge, const float *sigma, const float *eps,
                         const float *pos, float box, double *lj, double *dh);

/* Reference semantics over the first n_nc molecules only (the non-crowders):
 * the CPU branch of src/Replica.cpp:1062-1068. */
void orc_energy_nc_ref(int n_nc, const int *mol_off, const int *type,
                       const float *charge, const float *sigma, const float *eps,
                       const float *pos, float box, double *lj, double *dh);

/* Molecule-pair matrix in kcal/mol, mat[I*n_mol+J] = mat[J*n_mol+I],
 * diagonal 0, each entry with reference-semantics pair terms. */
void orc_energy_matrix_ref(int n_mol, const int *mol_off, const int *type,
                           const float *charge, const float *sigma, const float *eps,
                           const float *pos, float box, double *mat);

/* FP64 oracle of the same model: every operation in double (positions and
 * table values are the float inputs), Neumaier-compensated sums.  Pinned by
 * closed-form two-residue cases and agreement with orc_energy_ref. */
/* Repulsive crowding (src/definitions.h:18, src/Replica.cpp:416-483) for
 * every energy function here: pairs with a residue at or past
 * first_crowder_residue take (6/r)^12 below 15 A and no DH; negative: off. */
void orc_set_repulsive_crowding(int first_crowder_residue);

void orc_energy_fp64(int n_mol, const int *mol_off, const int *type,
                     const float *charge, const float *sigma, const float *eps,
                     const float *pos, float box, double *lj, double *dh);

/* ---------------------------------------------------------------------
 * Flexible linkers (the reference's FLEXIBLE_LINKS build).
 *
 * Topology of the whole system, in caller order (global residue indices):
 * lists are CSR per molecule unless stated.  Built by oracle/topology.py
 * (a restatement of src/Geometry.cpp:34-302).
 * --------------------------------------------------------------------- */
typedef struct {
    const int *mol_flexible;       /* M: 1 when the molecule has a segment or is all flexible */
    const int *tag;                /* N*4: chain UID, rigid-domain UID, segment-bond UID, resSeq */
    const int *bond_off, *bond;    /* M+1; 2 ints per bond (Graph::bonds order) */
    const int *angle_off, *angle;  /* M+1; 3 ints per angle */
    const int *tors_off, *tors;    /* M+1; 4 ints per torsion */
    const int *ind_off, *ind;      /* M+1; 2 ints per indirect-neighbour pair */
    const int *local_off, *local;  /* M+1; MC_local_residues */
    const int *crank_off, *crank;  /* M+1; MC_crankshaft_residues */
    const int *flexr_off, *flexr;  /* M+1; MC_flex_residues */
    const int *nb_off, *nb;        /* per flex-residue entry: its neighbours (Graph::neighbours order) */
    const int *br_off, *br;        /* per neighbour entry: the residues flex(ri, nb) rotates, ascending */
    const double *tor_v, *tor_d;   /* 20*20*4 each, [a][b][n-1]; phase in radians */
} orc_flex;

/* Bonded terms with the reference's float vectors and double potentials:
 * calculate_bond src/Potential.cpp:45-55 ((r - R0)^2, minimum image),
 * calculate_angle :57-71 (src/vector3f.cpp:137-143 for the angle),
 * calculate_torsion :73-98.  Pinned against oracle/_ref/ref_flex_driver
 * (the reference's own functions) and tests/testMolecule.cpp:287-303. */
double orc_bond_term(const float *pi, const float *pj, float box);
double orc_angle_term(const float *pi, const float *pj, const float *pk);
double orc_torsion_term(const float *pi, const float *pj, const float *pk, const float *pl, int type_j,
                        int type_k, const double *tor_v, const double *tor_d);

/* Internal energy of molecule m, reference semantics, out[5] = raw LJ,
 * raw DH, raw bond sum, angle product, torsion (src/Potential.h layout):
 *   mode 0  Molecule::E(true)  src/Molecule.cpp:636-699 (CPU exclusions)
 *   mode 1  Molecule::E(false) (minus the indirect neighbours, + bonded)
 *   mode 2  the GPU kernel's intra-molecular pairs only
 *           (src/CudaFunctions.cu:447-476: same rigid domain, same segment
 *           bond, |index difference| < 4 on one chain excluded)
 * Rigid molecules give (0, 0, 0, 1, 0).  Pinned against ref_flex_driver. */
void orc_internal_ref(const orc_flex *f, int m, const int *mol_off, const int *type, const float *charge,
                      const float *sigma, const float *eps, const float *pos, float box, int mode, double out[5]);

/* The same three modes with every operation in double (FP64 oracle). */
void orc_internal_fp64(const orc_flex *f, int m, const int *mol_off, const int *type, const float *charge,
                       const float *sigma, const float *eps, const float *pos, float box, int mode,
                       double out[5]);

/* Potential::total() of such a vector (src/Potential.cpp:203-210). */
double orc_internal_total(const double v[5]);

/* Flexible moves of one molecule, residues [0, n) in caller order, positions
 * pos/rel (n*3) and centre c[3] updated in place, then wrapped
 * (src/Molecule.cpp:242-252):
 *   orc_local_translate  translate(v, ri)             src/Molecule.cpp:274-301
 *   orc_crankshaft       crankshaft(angle, flip, ri)  :303-329
 *                        (rcos, rsin = cos, sin of angle/2)
 *   orc_flex_move        flex(axis, angle, ri, nb) on the given branch
 *                        (ascending residue indices)  :331-395
 * Pinned by the explicit coordinates of tests/testMolecule.cpp:146-283. */
void orc_local_translate(float *pos, float *rel, float c[3], int n, int ri, const float v[3], float box);
void orc_crankshaft(float *pos, float *rel, float c[3], int n, int ri, double rcos, double rsin, int flip,
                    float box);
void orc_flex_move(float *pos, float *rel, float c[3], int n, int ri, const double axis[3], double rcos,
                   double rsin, int n_branch, const int *branch, float box);

/* ---------------------------------------------------------------------
 * Rigid-body MC (device algorithm restated on the CPU).
 *
 * The device keeps replicas resident and proposes moves from counter-based
 * Philox streams, so this CPU restatement reproduces the same trajectory.
 * Moves follow src/Molecule.cpp:146-266,468-504 and src/Quaternion.cpp:12-112;
 * Metropolis follows src/Replica.cpp:385-410 (delta and exponent in FP64);
 * exchange follows src/Simulation.cpp:770-795 and src/Replica.cpp:188-210.
 * --------------------------------------------------------------------- */
typedef struct {
    int n_mol, n_res, n_rep;
    const int *mol_off;       /* n_mol+1 */
    const int *type;          /* n_res   */
    const float *charge, *sigma, *eps;
    float box;
    uint64_t seed;
    /* per-replica state, row-major by replica */
    float *center;            /* R*M*3 */
    float *rel;               /* R*N*3 */
    float *pos;               /* R*N*3 */
    double *temperature;      /* R */
    float *translate_step;    /* R */
    float *rotate_step;       /* R */
    double *energy;           /* R, kcal/mol */
    double *emat;             /* R*M*M, kcal/mol */
    int64_t *counters;        /* R*3: accept (dE<0), acceptA (Boltzmann), reject */
    int64_t *step;            /* R   MC steps taken by each replica */
    int *label;               /* R */
    int *rank_of;             /* R   temperature rank of replica r */
    int *rep_at_rank;         /* R   replica holding temperature rank k */
    int64_t *exchange_round;  /* 1 */
    int64_t *exchange_stats;  /* 2: tests, accepted swaps */
    const orc_flex *flex;     /* NULL: rigid molecules only */
} orc_remc;

/* Initialise positions (pos = rel + center), the molecule-pair matrix and
 * the per-replica totals. */
void orc_remc_init(orc_remc *s);

/* The proposal of step `step` of replica r, exactly as the device draws it.
 * Outputs: molecule, move kind (1 = rotate, 0 = translate), translation
 * vector (float) or unit axis (double), and the Metropolis uniform. */
void orc_mc_proposal(uint64_t seed, int r, int64_t step, int n_mol,
                     int *mol, int *is_rotate, double axis[3], double *u_accept);

/* The same draw with flexible molecules: kind 0 translate, 1 rotate,
 * 2 flex, 3 local (src/Molecule.cpp:451-466: equal weights over the four,
 * or over rotate/flex/local when the flexible molecule is the only one);
 * rigid molecules keep the Bernoulli(0.5) of orc_mc_proposal.  extra[4]
 * are the words that pick the residue, the neighbour, crankshaft vs
 * translate and the crankshaft direction. */
void orc_mc_proposal_flex(uint64_t seed, int r, int64_t step, int n_mol, const int *mol_flexible, int *mol,
                          int *kind, double axis[3], double *u_accept, uint32_t extra[4]);

/* Applies one rigid move to n residues (translation when is_rotate == 0):
 * the move code of orc_remc_mc_steps, exposed for the move tests.  rcos,
 * rsin = cos(angle/2), sin(angle/2). */
void orc_apply_move(int is_rotate, const double axis_raw[3], float tstep, double rcos, double rsin,
                    float box, const float c_in[3], float c_out[3], int n, const float *rel_in,
                    float *rel_out, float *pos_out);

/* The Metropolis decision of src/Replica.cpp:385-410 on an energy change
 * delta (kcal/mol) at temperature T (K) with the uniform u in [0, 1):
 * 0 downhill (delta < 0, the reference's `accept`), 1 accepted by the
 * Boltzmann test u < exp(-delta kcal / (Rgas T)) (`acceptA`), 2 rejected
 * (`reject`).  orc_metropolis takes an FP64 delta and exponent (the delta-E
 * loop and the device); orc_metropolis_ref the reference's float delta and
 * float exponent (the CPU baseline). */
int orc_metropolis(double delta, double temperature, double u);
int orc_metropolis_ref(float delta, float temperature, double u);

/* nsteps MC steps on every replica (delta-E rows against the matrix). */
void orc_remc_mc_steps(orc_remc *s, int nsteps);

/* One exchange round over adjacent temperature ranks. */
void orc_remc_exchange(orc_remc *s);

/* Lean FP64 sampler (rigid molecules, crowding off): the Markov chain of
 * orc_remc_mc_steps with an FP64 pair term written for speed, replicas over
 * threads; for CPU REMC runs long enough for statistical comparisons.
 * orc_remc_refresh_fast rebuilds the matrices and totals with the same pair
 * term; orc_energy_fast is replica r's total (kcal/mol). */
void orc_remc_refresh_fast(orc_remc *s);
void orc_remc_mc_steps_fast(orc_remc *s, int nsteps, int threads);
double orc_energy_fast(const orc_remc *s, int r);

/* ---------------------------------------------------------------------
 * CPU baseline: the reference's CPU REMC loop (src/Replica.cpp:350-414,
 * src/Simulation.cpp:996-1039): full Replica::E() recompute every MC step,
 * float Metropolis delta, restore on reject; replicas split over pthreads
 * as in src/Simulation.cpp:479-487.  Moves come from the same Philox
 * streams.  Returns wall seconds.
 * --------------------------------------------------------------------- */
double orc_remc_reference_run(orc_remc *s, int nsteps, int threads);

#ifdef __cplusplus
}
#endif
#endif
