/*
 * oracle.c -- CPU restatement of cgppd's REMC energy hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker and as the CPU baseline.  The
 * product (cgppd_amd/) never links or loads it.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, so every float
 * expression rounds the way the reference's x86-64 build rounds it).
 *
 * Parity status of each function is stated in oracle.h.
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* src/constants.h:8-26 and src/definitions.h:10 */
#define ORC_E_CHARGE 1.602176487e-19f
#define ORC_ANG 1e-10f
#define ORC_D_WATER 80.0f
#define ORC_RGAS 8.3144621f
#define ORC_T_ROOM 298.0f
#define ORC_KCAL 4184.0f
#define ORC_N_A 6.02214129e23f
#define ORC_EPS0 8.85418782e-12f
#define ORC_E0 (-2.27f)
#define ORC_LAMBDA 0.159f
#define ORC_XI 10.0f
#define ORC_EPS 1e-38f

float orc_rt_to_kcal(void) { return ORC_T_ROOM * ORC_RGAS / ORC_KCAL; }

float orc_dh_factor(void)
{
    /* float pow and float arithmetic, left to right, as the macro expands */
    return powf(ORC_E_CHARGE, 2.0f) * ORC_N_A /
           (4.0f * (float)3.14159265358979323846264338327950288L * ORC_EPS0 * ORC_KCAL * ORC_ANG * ORC_D_WATER);
}

float orc_r0_factor(void) { return powf(2.0f, 1.0f / 6.0f); }

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                        */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------ */
/* Reference-semantics pair term                                       */
/* ------------------------------------------------------------------ */

/* src/Residue.cpp:62-76: float difference, float minimum image with
 * round(), sqrtf; src/Replica.cpp:479 adds EPS in double. */
static inline double ref_distance(const float *a, const float *b, float box)
{
    float X = a[0] - b[0];
    float Y = a[1] - b[1];
    float Z = a[2] - b[2];
    if (box != 0.0f) {
        X = X - box * roundf(X / box);
        Y = Y - box * roundf(Y / box);
        Z = Z - box * roundf(Z / box);
    }
    return (double)sqrtf(X * X + Y * Y + Z * Z) + (double)ORC_EPS;
}

/* src/Potential.cpp:5-35 */
static inline double ref_lj(int ti, int tj, const float *sigma, const float *eps, double r, float r0f)
{
    double Eij = (double)(ORC_LAMBDA * (eps[ti * 20 + tj] - ORC_E0));
    double sij = (double)(0.5f * (sigma[ti] + sigma[tj]));
    double t = pow(sij / r, 6.0);
    double lj = -4.0 * Eij * t * (t - 1.0);
    if (Eij > 0.0 && r < sij * (double)r0f)
        lj = -lj + 2.0 * Eij;
    return lj;
}

/* src/Potential.cpp:37-42 */
static inline double ref_dh(int ti, int tj, const float *charge, double r)
{
    float e = expf((float)(-r / (double)ORC_XI));
    return (double)(charge[ti] * charge[tj] * e) / r;
}

/* Repulsive crowding (the reference's compile-time REPULSIVE_CROWDING,
 * src/definitions.h:18; src/Replica.cpp:416-483): a pair with a residue of
 * a crowder molecule contributes crowderPairPotential(r) = (6/r)^12 (float
 * powf) to LJ for r < const_repulsive_cutoff = 15 A (src/definitions.h:38)
 * and no Debye-Hueckel term.  Crowders are the molecules from the first
 * crowder residue on; off (the reference's default) when negative. */
static int orc_crowd_res = 0x7fffffff;
void orc_set_repulsive_crowding(int first_crowder_residue)
{
    orc_crowd_res = first_crowder_residue < 0 ? 0x7fffffff : first_crowder_residue;
}
#define ORC_REPULSIVE_CUTOFF 15.0f

static void ref_block(int i0, int i1, int j0, int j1, const int *type, const float *charge,
                      const float *sigma, const float *eps, const float *pos, float box,
                      double *lj, double *dh)
{
    const float r0f = orc_r0_factor();
    double sl = *lj, sd = *dh;
    for (int i = i0; i < i1; ++i) {
        for (int j = j0; j < j1; ++j) {
            double r = ref_distance(pos + 3 * i, pos + 3 * j, box);
            if (i >= orc_crowd_res || j >= orc_crowd_res) {
                if (r < ORC_REPULSIVE_CUTOFF) sl += (double)powf(6.0f / (float)r, 12.0f);
                continue;
            }
            sl += ref_lj(type[i], type[j], sigma, eps, r, r0f);
            sd += ref_dh(type[i], type[j], charge, r);
        }
    }
    *lj = sl;
    *dh = sd;
}

/* src/Replica.cpp:421-500 (non-crowding branch, LJ on) */
void orc_energy_ref(int n_mol, const int *mol_off, const int *type, const float *charge,
                    const float *sigma, const float *eps, const float *pos, float box,
                    double *lj, double *dh)
{
    double sl = 0.0, sd = 0.0;
    for (int I = 0; I < n_mol; ++I)
        for (int J = I + 1; J < n_mol; ++J)
            ref_block(mol_off[I], mol_off[I + 1], mol_off[J], mol_off[J + 1], type, charge, sigma,
                      eps, pos, box, &sl, &sd);
    *lj = sl;
    *dh = sd;
}

/* src/Replica.cpp:502-528 */
void orc_energy_pair_ref(const int *mol_off, int a, int b, const int *type, const float *charge,
                         const float *sigma, const float *eps, const float *pos, float box,
                         double *lj, double *dh)
{
    double sl = 0.0, sd = 0.0;
    ref_block(mol_off[a], mol_off[a + 1], mol_off[b], mol_off[b + 1], type, charge, sigma, eps,
              pos, box, &sl, &sd);
    *lj = sl;
    *dh = sd;
}

/* src/Replica.cpp:1062-1068: sum of E(a,b) over non-crowder pairs.  The
 * reference adds totals per pair; the raw sums here are the same pairs. */
void orc_energy_nc_ref(int n_nc, const int *mol_off, const int *type, const float *charge,
                       const float *sigma, const float *eps, const float *pos, float box,
                       double *lj, double *dh)
{
    orc_energy_ref(n_nc, mol_off, type, charge, sigma, eps, pos, box, lj, dh);
}

static inline double pair_total(double lj, double dh)
{
    /* src/Potential.cpp:219 */
    return dh * (double)orc_dh_factor() + lj * (double)orc_rt_to_kcal();
}

void orc_energy_matrix_ref(int n_mol, const int *mol_off, const int *type, const float *charge,
                           const float *sigma, const float *eps, const float *pos, float box,
                           double *mat)
{
    for (int I = 0; I < n_mol; ++I) {
        mat[I * n_mol + I] = 0.0;
        for (int J = I + 1; J < n_mol; ++J) {
            double lj, dh;
            orc_energy_pair_ref(mol_off, I, J, type, charge, sigma, eps, pos, box, &lj, &dh);
            mat[I * n_mol + J] = mat[J * n_mol + I] = pair_total(lj, dh);
        }
    }
}

/* ------------------------------------------------------------------ */
/* FP64 oracle                                                          */
/* ------------------------------------------------------------------ */
typedef struct { double s, c; } nsum;
static inline void nsum_add(nsum *a, double v)
{
    double t = a->s + v;
    if (fabs(a->s) >= fabs(v)) a->c += (a->s - t) + v;
    else a->c += (v - t) + a->s;
    a->s = t;
}

void orc_energy_fp64(int n_mol, const int *mol_off, const int *type, const float *charge,
                     const float *sigma, const float *eps, const float *pos, float box,
                     double *lj, double *dh)
{
    const double L = (double)box;
    const double r0 = pow(2.0, 1.0 / 6.0);
    nsum sl = {0, 0}, sd = {0, 0};
    for (int I = 0; I < n_mol; ++I)
        for (int J = I + 1; J < n_mol; ++J)
            for (int i = mol_off[I]; i < mol_off[I + 1]; ++i)
                for (int j = mol_off[J]; j < mol_off[J + 1]; ++j) {
                    double d[3];
                    for (int k = 0; k < 3; ++k) {
                        d[k] = (double)pos[3 * i + k] - (double)pos[3 * j + k];
                        if (L != 0.0) d[k] -= L * round(d[k] / L);
                    }
                    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + 1e-38;
                    if (i >= orc_crowd_res || j >= orc_crowd_res) {
                        if (r < (double)ORC_REPULSIVE_CUTOFF) nsum_add(&sl, pow(6.0 / r, 12.0));
                        continue;
                    }
                    int ti = type[i], tj = type[j];
                    double Eij = (double)ORC_LAMBDA * ((double)eps[ti * 20 + tj] - (double)ORC_E0);
                    double sij = 0.5 * ((double)sigma[ti] + (double)sigma[tj]);
                    double t = pow(sij / r, 6.0);
                    double u = -4.0 * Eij * t * (t - 1.0);
                    if (Eij > 0.0 && r < sij * r0) u = -u + 2.0 * Eij;
                    nsum_add(&sl, u);
                    nsum_add(&sd, (double)charge[ti] * (double)charge[tj] * exp(-r / (double)ORC_XI) / r);
                }
    *lj = sl.s + sl.c;
    *dh = sd.s + sd.c;
}

/* ------------------------------------------------------------------ */
/* Flexible linkers: bonded terms and intra-molecular pairs            */
/* ------------------------------------------------------------------ */
/* src/constants.h:34-46 */
#define ORC_K_SPRING 378
#define ORC_R0 3.81f
#define ORC_GAMMA_ANGLE 0.1f
#define ORC_GAMMA_ANGLE_RECIP 10.0f
#define ORC_EPSILON_ALPHA 4.3f
#define ORC_THETA_ALPHA 1.6f
#define ORC_THETA_BETA 2.27f
#define ORC_K_ALPHA 106.4f
#define ORC_K_BETA 26.3f

/* src/vector3f.h:59-61: float magnitude */
static inline float vmag(const float v[3]) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static inline float vdot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void vcross(const float a[3], const float b[3], float out[3])
{
    /* src/vector3f.cpp:152-155 */
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

double orc_bond_term(const float *pi, const float *pj, float box)
{
    /* src/Potential.cpp:45-55: rj.distance(ri) without EPS */
    float X = pj[0] - pi[0], Y = pj[1] - pi[1], Z = pj[2] - pi[2];
    if (box != 0.0f) {
        X = X - box * roundf(X / box);
        Y = Y - box * roundf(Y / box);
        Z = Z - box * roundf(Z / box);
    }
    double r = (double)sqrtf(X * X + Y * Y + Z * Z);
    return (r - (double)ORC_R0) * (r - (double)ORC_R0);
}

double orc_angle_term(const float *pi, const float *pj, const float *pk)
{
    /* src/Potential.cpp:57-71 with Vector3f::angle (src/vector3f.cpp:137-143) */
    float ab[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
    float cb[3] = {pk[0] - pj[0], pk[1] - pj[1], pk[2] - pj[2]};
    double cmag = (double)(vmag(cb) * vmag(ab));
    float thf;
    if (cmag < 1e-13 * 1e-13) thf = 0.0f;
    else thf = (float)acos((double)vdot(ab, cb) / cmag);
    double theta = (double)thf;
    double a = theta - (double)ORC_THETA_ALPHA, b = theta - (double)ORC_THETA_BETA;
    return exp(-(double)ORC_GAMMA_ANGLE * ((double)ORC_K_ALPHA * a * a + (double)ORC_EPSILON_ALPHA)) +
           exp(-(double)ORC_GAMMA_ANGLE * ((double)ORC_K_BETA * b * b));
}

double orc_torsion_term(const float *pi, const float *pj, const float *pk, const float *pl, int type_j,
                        int type_k, const double *tor_v, const double *tor_d)
{
    /* src/Potential.cpp:73-98: float vectors, float atan2 (the arguments
     * are floats), double cosines */
    float b1[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
    float b2[3] = {pk[0] - pj[0], pk[1] - pj[1], pk[2] - pj[2]};
    float b3[3] = {pl[0] - pk[0], pl[1] - pk[1], pl[2] - pk[2]};
    float b2xb3[3], b1xb2[3];
    vcross(b2, b3, b2xb3);
    vcross(b1, b2, b1xb2);
    double phi = (double)atan2f(vmag(b2) * vdot(b1, b2xb3), vdot(b1xb2, b2xb3));
    const double *V = tor_v + (type_j * 20 + type_k) * 4;
    const double *D = tor_d + (type_j * 20 + type_k) * 4;
    return (1 + cos(phi - D[0])) * V[0] + (1 + cos(2 * phi - D[1])) * V[1] + (1 + cos(3 * phi - D[2])) * V[2] +
           (1 + cos(4 * phi - D[3])) * V[3];
}

static int is_indirect(const orc_flex *f, int m, int i, int j)
{
    for (int k = f->ind_off[m]; k < f->ind_off[m + 1]; ++k)
        if (f->ind[2 * k] == i && f->ind[2 * k + 1] == j) return 1;
    return 0;
}

/* src/Residue.cpp:116-129 (CPU) and src/CudaFunctions.cu:447-476 (GPU) */
static int excluded(const orc_flex *f, int m, int i, int j, int gpu_rule)
{
    const int *ti = f->tag + 4 * i, *tj = f->tag + 4 * j;
    if (ti[1] && ti[1] == tj[1]) return 1;
    if (ti[2] && ti[2] == tj[2]) return 1;
    if (gpu_rule) return ti[0] == tj[0] && (j - i) < 4;
    if (ti[0] == tj[0] && abs(ti[3] - tj[3]) < 4) return 1;
    return is_indirect(f, m, i, j);
}

void orc_internal_ref(const orc_flex *f, int m, const int *mol_off, const int *type, const float *charge,
                      const float *sigma, const float *eps, const float *pos, float box, int mode, double out[5])
{
    out[0] = out[1] = out[2] = 0.0;
    out[3] = 1.0;
    out[4] = 0.0;
    if (!f->mol_flexible[m]) return;
    const float r0f = orc_r0_factor();
    const int i0 = mol_off[m], i1 = mol_off[m + 1];
    if (mode == 0 || mode == 2) {
        for (int i = i0; i < i1; ++i)
            for (int j = i + 1; j < i1; ++j) {
                if (excluded(f, m, i, j, mode == 2)) continue;
                double r = ref_distance(pos + 3 * i, pos + 3 * j, box);
                out[0] += ref_lj(type[i], type[j], sigma, eps, r, r0f);
                out[1] += ref_dh(type[i], type[j], charge, r);
            }
    }
    if (mode == 1) {
        for (int k = f->ind_off[m]; k < f->ind_off[m + 1]; ++k) {
            const int i = f->ind[2 * k], j = f->ind[2 * k + 1];
            double r = ref_distance(pos + 3 * i, pos + 3 * j, box);
            out[0] += -ref_lj(type[i], type[j], sigma, eps, r, r0f);
            out[1] += -ref_dh(type[i], type[j], charge, r);
        }
    }
    if (mode == 2) return;
    for (int k = f->bond_off[m]; k < f->bond_off[m + 1]; ++k)
        out[2] += orc_bond_term(pos + 3 * f->bond[2 * k], pos + 3 * f->bond[2 * k + 1], box);
    for (int k = f->angle_off[m]; k < f->angle_off[m + 1]; ++k) {
        const int *a = f->angle + 3 * k;
        out[3] *= orc_angle_term(pos + 3 * a[0], pos + 3 * a[1], pos + 3 * a[2]);
    }
    for (int k = f->tors_off[m]; k < f->tors_off[m + 1]; ++k) {
        const int *t = f->tors + 4 * k;
        out[4] += orc_torsion_term(pos + 3 * t[0], pos + 3 * t[1], pos + 3 * t[2], pos + 3 * t[3], type[t[1]],
                                   type[t[2]], f->tor_v, f->tor_d);
    }
}

/* FP64 pair term and bonded terms */
static void fp64_pair(const float *a, const float *b, int ti, int tj, const float *charge, const float *sigma,
                      const float *eps, float box, double *lj, double *dh)
{
    const double L = (double)box;
    double d[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = (double)a[k] - (double)b[k];
        if (L != 0.0) d[k] -= L * round(d[k] / L);
    }
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + 1e-38;
    double Eij = (double)ORC_LAMBDA * ((double)eps[ti * 20 + tj] - (double)ORC_E0);
    double sij = 0.5 * ((double)sigma[ti] + (double)sigma[tj]);
    double t = pow(sij / r, 6.0);
    double u = -4.0 * Eij * t * (t - 1.0);
    if (Eij > 0.0 && r < sij * pow(2.0, 1.0 / 6.0)) u = -u + 2.0 * Eij;
    *lj = u;
    *dh = (double)charge[ti] * (double)charge[tj] * exp(-r / (double)ORC_XI) / r;
}

void orc_internal_fp64(const orc_flex *f, int m, const int *mol_off, const int *type, const float *charge,
                       const float *sigma, const float *eps, const float *pos, float box, int mode,
                       double out[5])
{
    out[0] = out[1] = out[2] = 0.0;
    out[3] = 1.0;
    out[4] = 0.0;
    if (!f->mol_flexible[m]) return;
    const int i0 = mol_off[m], i1 = mol_off[m + 1];
    nsum sl = {0, 0}, sd = {0, 0};
    if (mode == 0 || mode == 2) {
        for (int i = i0; i < i1; ++i)
            for (int j = i + 1; j < i1; ++j) {
                if (excluded(f, m, i, j, mode == 2)) continue;
                double lj, dh;
                fp64_pair(pos + 3 * i, pos + 3 * j, type[i], type[j], charge, sigma, eps, box, &lj, &dh);
                nsum_add(&sl, lj);
                nsum_add(&sd, dh);
            }
    }
    if (mode == 1) {
        for (int k = f->ind_off[m]; k < f->ind_off[m + 1]; ++k) {
            const int i = f->ind[2 * k], j = f->ind[2 * k + 1];
            double lj, dh;
            fp64_pair(pos + 3 * i, pos + 3 * j, type[i], type[j], charge, sigma, eps, box, &lj, &dh);
            nsum_add(&sl, -lj);
            nsum_add(&sd, -dh);
        }
    }
    out[0] = sl.s + sl.c;
    out[1] = sd.s + sd.c;
    if (mode == 2) return;
    const double L = (double)box;
    nsum sb = {0, 0}, st = {0, 0};
    double logang = 0.0;
    for (int k = f->bond_off[m]; k < f->bond_off[m + 1]; ++k) {
        const float *a = pos + 3 * f->bond[2 * k], *b = pos + 3 * f->bond[2 * k + 1];
        double d[3];
        for (int q = 0; q < 3; ++q) {
            d[q] = (double)b[q] - (double)a[q];
            if (L != 0.0) d[q] -= L * round(d[q] / L);
        }
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        nsum_add(&sb, (r - 3.81) * (r - 3.81));
    }
    for (int k = f->angle_off[m]; k < f->angle_off[m + 1]; ++k) {
        const int *t = f->angle + 3 * k;
        double ab[3], cb[3];
        for (int q = 0; q < 3; ++q) {
            ab[q] = (double)pos[3 * t[0] + q] - (double)pos[3 * t[1] + q];
            cb[q] = (double)pos[3 * t[2] + q] - (double)pos[3 * t[1] + q];
        }
        double c = (ab[0] * cb[0] + ab[1] * cb[1] + ab[2] * cb[2]) /
                   (sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]) * sqrt(cb[0] * cb[0] + cb[1] * cb[1] + cb[2] * cb[2]));
        double th = acos(c);
        double a = th - 1.6, b = th - 2.27;
        logang += log(exp(-0.1 * (106.4 * a * a + 4.3)) + exp(-0.1 * (26.3 * b * b)));
    }
    for (int k = f->tors_off[m]; k < f->tors_off[m + 1]; ++k) {
        const int *t = f->tors + 4 * k;
        double b1[3], b2[3], b3[3], x23[3], x12[3];
        for (int q = 0; q < 3; ++q) {
            b1[q] = (double)pos[3 * t[1] + q] - (double)pos[3 * t[0] + q];
            b2[q] = (double)pos[3 * t[2] + q] - (double)pos[3 * t[1] + q];
            b3[q] = (double)pos[3 * t[3] + q] - (double)pos[3 * t[2] + q];
        }
        x23[0] = b2[1] * b3[2] - b2[2] * b3[1];
        x23[1] = b2[2] * b3[0] - b2[0] * b3[2];
        x23[2] = b2[0] * b3[1] - b2[1] * b3[0];
        x12[0] = b1[1] * b2[2] - b1[2] * b2[1];
        x12[1] = b1[2] * b2[0] - b1[0] * b2[2];
        x12[2] = b1[0] * b2[1] - b1[1] * b2[0];
        double m2 = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
        double phi = atan2(m2 * (b1[0] * x23[0] + b1[1] * x23[1] + b1[2] * x23[2]),
                           x12[0] * x23[0] + x12[1] * x23[1] + x12[2] * x23[2]);
        const double *V = f->tor_v + (type[t[1]] * 20 + type[t[2]]) * 4;
        const double *D = f->tor_d + (type[t[1]] * 20 + type[t[2]]) * 4;
        for (int n = 0; n < 4; ++n) nsum_add(&st, (1.0 + cos((n + 1) * phi - D[n])) * V[n]);
    }
    out[2] = sb.s + sb.c;
    out[3] = exp(logang);
    out[4] = st.s + st.c;
}

double orc_internal_total(const double v[5])
{
    /* src/Potential.cpp:203-210 */
    return v[1] * (double)orc_dh_factor() + v[0] * (double)orc_rt_to_kcal() + v[2] * 0.5 * ORC_K_SPRING +
           log(v[3]) * -(double)ORC_GAMMA_ANGLE_RECIP + v[4];
}

/* src/Quaternion.cpp:95-112 */
static inline void qrot(const double q[4], const float v[3], float out[3])
{
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double w2 = w * w, x2 = x * x, y2 = y * y, z2 = z * z;
    const double wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
    const double vx = v[0], vy = v[1], vz = v[2];
    out[0] = (float)(vx * (w2 + x2 - y2 - z2) + (vy * (xy - wz) + vz * (wy + xz)) * 2.0);
    out[1] = (float)((vx * (wz + xy) + vz * (yz - wx)) * 2.0 + vy * (w2 - x2 + y2 - z2));
    out[2] = (float)((vx * (xz - wy) + vy * (yz + wx)) * 2.0 + vz * (w2 - x2 - y2 + z2));
}

/* src/Molecule.cpp:242-252 wrap_if_necessary with wrapped_centre :234-241 */
static void wrap_molecule(float *pos, const float *rel, float c[3], int n, float box)
{
    float w[3] = {fmodf(c[0], box), fmodf(c[1], box), fmodf(c[2], box)};
    float d[3] = {w[0] - c[0], w[1] - c[1], w[2] - c[2]};
    if (vmag(d) != 0.0f) {
        c[0] = w[0]; c[1] = w[1]; c[2] = w[2];
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) pos[3 * i + k] = rel[3 * i + k] + c[k];
    }
}

void orc_local_translate(float *pos, float *rel, float c[3], int n, int ri, const float v[3], float box)
{
    const float nf = (float)n;
    float vf[3], nc[3];
    for (int k = 0; k < 3; ++k) {
        vf[k] = v[k] / nf;
        nc[k] = c[k] + vf[k];
    }
    for (int k = 0; k < 3; ++k) {
        pos[3 * ri + k] = pos[3 * ri + k] + v[k];
        rel[3 * ri + k] = rel[3 * ri + k] + v[k];
    }
    c[0] = nc[0]; c[1] = nc[1]; c[2] = nc[2];
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) rel[3 * i + k] = rel[3 * i + k] - vf[k];
    wrap_molecule(pos, rel, c, n, box);
}

void orc_crankshaft(float *pos, float *rel, float c[3], int n, int ri, double rcos, double rsin, int flip,
                    float box)
{
    /* axis from the neighbouring residues, float difference, double normalise */
    double ax[3];
    for (int k = 0; k < 3; ++k) ax[k] = (double)(pos[3 * (ri + 1) + k] - pos[3 * (ri - 1) + k]);
    double m = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    ax[0] /= m; ax[1] /= m; ax[2] /= m;
    const double s = flip ? -rsin : rsin;
    double q[4] = {rcos, s * ax[0], s * ax[1], s * ax[2]};
    double qm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    q[0] /= qm; q[1] /= qm; q[2] /= qm; q[3] /= qm;
    float rp[3], rr[3], v[3];
    for (int k = 0; k < 3; ++k) rp[k] = pos[3 * ri + k] - pos[3 * (ri - 1) + k];
    qrot(q, rp, rr);
    for (int k = 0; k < 3; ++k) v[k] = rr[k] + pos[3 * (ri - 1) + k] - pos[3 * ri + k];
    orc_local_translate(pos, rel, c, n, ri, v, box);
}

void orc_flex_move(float *pos, float *rel, float c[3], int n, int ri, const double axis[3], double rcos,
                   double rsin, int n_branch, const int *branch, float box)
{
    /* Quaternion(angle, axis) without normalisation (src/Molecule.cpp:333) */
    const double q[4] = {rcos, rsin * axis[0], rsin * axis[1], rsin * axis[2]};
    float acc[3] = {0.0f, 0.0f, 0.0f};
    float *diff = (float *)malloc(sizeof(float) * 3 * (n_branch > 0 ? n_branch : 1));
    const float pr[3] = {pos[3 * ri], pos[3 * ri + 1], pos[3 * ri + 2]};
    for (int b = 0; b < n_branch; ++b) {
        const int i = branch[b];
        float rp[3], rr[3];
        for (int k = 0; k < 3; ++k) rp[k] = pos[3 * i + k] - pr[k];
        qrot(q, rp, rr);
        for (int k = 0; k < 3; ++k) {
            diff[3 * b + k] = rr[k] + pr[k] - pos[3 * i + k];
            acc[k] = acc[k] + diff[3 * b + k];
        }
    }
    const float nf = (float)n;
    float vf[3], nc[3];
    for (int k = 0; k < 3; ++k) {
        vf[k] = acc[k] / nf;
        nc[k] = c[k] + vf[k];
    }
    for (int b = 0; b < n_branch; ++b) {
        const int i = branch[b];
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] = pos[3 * i + k] + diff[3 * b + k];
            rel[3 * i + k] = rel[3 * i + k] + diff[3 * b + k];
        }
    }
    free(diff);
    c[0] = nc[0]; c[1] = nc[1]; c[2] = nc[2];
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) rel[3 * i + k] = rel[3 * i + k] - vf[k];
    wrap_molecule(pos, rel, c, n, box);
}

/* ------------------------------------------------------------------ */
/* Rigid-body moves                                                    */
/* ------------------------------------------------------------------ */
#define ORC_TAG_MC 0x4D430000u
#define ORC_TAG_EX 0x45580000u

static inline double u32_to_unit(uint32_t w) { return (double)w * (1.0 / 4294967296.0); }
static inline double u53(uint32_t hi, uint32_t lo)
{
    return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

void orc_mc_proposal(uint64_t seed, int r, int64_t step, int n_mol, int *mol, int *is_rotate,
                     double axis[3], double *u_accept)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t c0[4] = {(uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)r, ORC_TAG_MC};
    uint32_t c1[4] = {(uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)r, ORC_TAG_MC | 1u};
    uint32_t w[4], v[4];
    orc_philox4x32_10(c0, key, w);
    orc_philox4x32_10(c1, key, v);
    /* src/Replica.cpp:355 picks the molecule uniformly */
    *mol = (int)(((uint64_t)w[0] * (uint64_t)n_mol) >> 32);
    /* src/Molecule.cpp:461: Bernoulli(0.5), 1 = MC_ROTATE */
    *is_rotate = w[1] < 0x80000000u ? 1 : 0;
    /* src/Molecule.cpp:213,218: (u-0.5, 0.5-u, u-0.5) before normalising */
    axis[0] = u32_to_unit(v[0]) - 0.5;
    axis[1] = 0.5 - u32_to_unit(v[1]);
    axis[2] = u32_to_unit(v[2]) - 0.5;
    *u_accept = u53(w[2], w[3]);
}

static inline int mulhi_n(uint32_t w, int n) { return (int)(((uint64_t)w * (uint64_t)n) >> 32); }

void orc_mc_proposal_flex(uint64_t seed, int r, int64_t step, int n_mol, const int *mol_flexible, int *mol,
                          int *kind, double axis[3], double *u_accept, uint32_t extra[4])
{
    int rot;
    orc_mc_proposal(seed, r, step, n_mol, mol, &rot, axis, u_accept);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t c0[4] = {(uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)r, ORC_TAG_MC};
    uint32_t c2[4] = {(uint32_t)step, (uint32_t)((uint64_t)step >> 32), (uint32_t)r, ORC_TAG_MC | 2u};
    uint32_t w[4];
    orc_philox4x32_10(c0, key, w);
    orc_philox4x32_10(c2, key, extra);
    if (mol_flexible && mol_flexible[*mol]) {
        /* src/Molecule.cpp:451-466 with the equal weights of src/Replica.cpp:309-334 */
        *kind = n_mol > 1 ? mulhi_n(w[1], 4) : 1 + mulhi_n(w[1], 3);
    } else {
        *kind = rot;
    }
}

/* Apply the proposal to molecule residues [i0,i1) of one replica; writes
 * the new centre, relative and absolute positions into the output arrays
 * (which may alias the inputs). */
static void apply_move(int is_rotate, const double axis_raw[3], float tstep, double rcos,
                       double rsin, float box, const float *c_in, float *c_out, int i0, int i1,
                       const float *rel_in, float *rel_out, float *pos_out)
{
    if (!is_rotate) {
        /* src/Molecule.cpp:262-266, 210-214: float vector, float normalise */
        float n[3] = {(float)axis_raw[0], (float)axis_raw[1], (float)axis_raw[2]};
        float mag = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (mag != 0.0f) {
            n[0] = n[0] / mag; n[1] = n[1] / mag; n[2] = n[2] / mag;
        }
        float c[3];
        for (int k = 0; k < 3; ++k) {
            float v = (float)((double)n[k] * (double)tstep);
            /* src/Molecule.cpp:146-159 + wrapped_centre :234-241 (fmod) */
            c[k] = fmodf(c_in[k] + v, box);
        }
        for (int k = 0; k < 3; ++k) c_out[k] = c[k];
        for (int i = i0; i < i1; ++i)
            for (int k = 0; k < 3; ++k) {
                rel_out[3 * i + k] = rel_in[3 * i + k];
                /* src/Molecule.cpp:183-191 */
                pos_out[3 * i + k] = rel_in[3 * i + k] + c[k];
            }
    } else {
        /* src/Molecule.cpp:216-221, 193-208; src/Quaternion.cpp:12-38,95-112 */
        double ax = axis_raw[0], ay = axis_raw[1], az = axis_raw[2];
        double m = sqrt(ax * ax + ay * ay + az * az);
        ax /= m; ay /= m; az /= m;
        double w = rcos, x = rsin * ax, y = rsin * ay, z = rsin * az;
        double qm = sqrt(w * w + x * x + y * y + z * z);
        w /= qm; x /= qm; y /= qm; z /= qm;
        double w2 = w * w, x2 = x * x, y2 = y * y, z2 = z * z;
        double wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
        float c[3] = {c_in[0], c_in[1], c_in[2]};
        for (int k = 0; k < 3; ++k) c_out[k] = c[k];
        for (int i = i0; i < i1; ++i) {
            double vx = rel_in[3 * i], vy = rel_in[3 * i + 1], vz = rel_in[3 * i + 2];
            float nx = (float)(vx * (w2 + x2 - y2 - z2) + (vy * (xy - wz) + vz * (wy + xz)) * 2.0);
            float ny = (float)((vx * (wz + xy) + vz * (yz - wx)) * 2.0 + vy * (w2 - x2 + y2 - z2));
            float nz = (float)((vx * (xz - wy) + vy * (yz + wx)) * 2.0 + vz * (w2 - x2 - y2 + z2));
            rel_out[3 * i] = nx; rel_out[3 * i + 1] = ny; rel_out[3 * i + 2] = nz;
            /* src/Residue.cpp:93-99 */
            pos_out[3 * i] = c[0] + nx;
            pos_out[3 * i + 1] = c[1] + ny;
            pos_out[3 * i + 2] = c[2] + nz;
        }
    }
}

void orc_apply_move(int is_rotate, const double axis_raw[3], float tstep, double rcos, double rsin,
                    float box, const float c_in[3], float c_out[3], int n, const float *rel_in,
                    float *rel_out, float *pos_out)
{
    apply_move(is_rotate, axis_raw, tstep, rcos, rsin, box, c_in, c_out, 0, n, rel_in, rel_out, pos_out);
}

/* The MC diagonal of a flexible molecule: the GPU kernel's intra-molecular
 * pairs plus Molecule::E(false) (src/Replica.cpp:362-371), kcal/mol. */
static double mc_internal(const orc_remc *s, int m, const float *pos)
{
    double a[5], b[5];
    orc_internal_ref(s->flex, m, s->mol_off, s->type, s->charge, s->sigma, s->eps, pos, s->box, 2, a);
    orc_internal_ref(s->flex, m, s->mol_off, s->type, s->charge, s->sigma, s->eps, pos, s->box, 1, b);
    double v[5] = {a[0] + b[0], a[1] + b[1], b[2], b[3], b[4]};
    return orc_internal_total(v);
}

/* Applies the proposal (kind, axis, extra) to molecule m of one replica:
 * pos/rel are the replica's arrays, c the molecule's centre (in place). */
static void apply_proposal(const orc_remc *s, int r, int m, int kind, const double axis_raw[3],
                           const uint32_t extra[4], double rcos, double rsin, float *pos, float *rel, float *c)
{
    const int i0 = s->mol_off[m], i1 = s->mol_off[m + 1], n = i1 - i0;
    if (kind <= 1) {
        apply_move(kind, axis_raw, s->translate_step[r], rcos, rsin, s->box, c, c, i0, i1, rel, rel, pos);
        return;
    }
    const orc_flex *f = s->flex;
    float *mp = pos + 3 * i0, *mr = rel + 3 * i0;
    if (kind == 2) {
        /* src/Molecule.cpp:397-409 */
        const int nf = f->flexr_off[m + 1] - f->flexr_off[m];
        if (nf == 0) return;
        const int e = f->flexr_off[m] + mulhi_n(extra[0], nf);
        const int deg = f->nb_off[e + 1] - f->nb_off[e];
        const int k = f->nb_off[e] + mulhi_n(extra[1], deg);
        double ax = axis_raw[0], ay = axis_raw[1], az = axis_raw[2];
        double mag = sqrt(ax * ax + ay * ay + az * az);
        const double unit[3] = {ax / mag, ay / mag, az / mag};
        const int nb = f->br_off[k + 1] - f->br_off[k];
        int *loc = (int *)malloc(sizeof(int) * (nb > 0 ? nb : 1));
        for (int b = 0; b < nb; ++b) loc[b] = f->br[f->br_off[k] + b] - i0;
        orc_flex_move(mp, mr, c, n, f->flexr[e] - i0, unit, rcos, rsin, nb, loc, s->box);
        free(loc);
        return;
    }
    /* kind 3, src/Molecule.cpp:415-448: one local move */
    if (extra[2] < 0x80000000u) {
        const int nc = f->crank_off[m + 1] - f->crank_off[m];
        if (nc == 0) return;
        const int ri = f->crank[f->crank_off[m] + mulhi_n(extra[0], nc)] - i0;
        orc_crankshaft(mp, mr, c, n, ri, rcos, rsin, extra[3] < 0x80000000u, s->box);
    } else {
        const int nl = f->local_off[m + 1] - f->local_off[m];
        if (nl == 0) return;
        const int ri = f->local[f->local_off[m] + mulhi_n(extra[0], nl)] - i0;
        float v3[3] = {(float)axis_raw[0], (float)axis_raw[1], (float)axis_raw[2]};
        float mag = sqrtf(v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2]);
        if (mag != 0.0f) {
            v3[0] = v3[0] / mag; v3[1] = v3[1] / mag; v3[2] = v3[2] / mag;
        }
        const double sc = (double)0.1f * (double)s->translate_step[r];
        float v[3];
        for (int q = 0; q < 3; ++q) v[q] = (float)((double)v3[q] * sc);
        orc_local_translate(mp, mr, c, n, ri, v, s->box);
    }
}

void orc_remc_init(orc_remc *s)
{
    const int M = s->n_mol, N = s->n_res;
    for (int r = 0; r < s->n_rep; ++r) {
        float *pos = s->pos + (size_t)r * N * 3;
        const float *rel = s->rel + (size_t)r * N * 3;
        const float *c = s->center + (size_t)r * M * 3;
        for (int I = 0; I < M; ++I)
            for (int i = s->mol_off[I]; i < s->mol_off[I + 1]; ++i)
                for (int k = 0; k < 3; ++k) pos[3 * i + k] = rel[3 * i + k] + c[3 * I + k];
        double *mat = s->emat + (size_t)r * M * M;
        orc_energy_matrix_ref(M, s->mol_off, s->type, s->charge, s->sigma, s->eps, pos, s->box, mat);
        if (s->flex)
            for (int I = 0; I < M; ++I)
                if (s->flex->mol_flexible[I]) mat[I * M + I] = mc_internal(s, I, pos);
        double e = 0.0;
        for (int I = 0; I < M; ++I)
            for (int J = I; J < M; ++J) e += mat[I * M + J];
        s->energy[r] = e;
    }
}

/* src/Replica.cpp:385-410: downhill moves are accepted, uphill ones when
 * the uniform falls below the Boltzmann factor of delta converted from
 * kcal/mol to J/mol */
int orc_metropolis(double delta, double temperature, double u)
{
    if (delta < 0.0) return 0;
    if (u < exp(-delta * 4184.0 / ((double)ORC_RGAS * temperature))) return 1;
    return 2;
}

/* the same with the reference's float delta and float exponent
 * (src/Replica.cpp:385, `float delta = newPotential - potential`) */
int orc_metropolis_ref(float delta, float temperature, double u)
{
    if (delta < 0.0f) return 0;
    if (u < (double)expf(-(delta * ORC_KCAL) / (ORC_RGAS * temperature))) return 1;
    return 2;
}

static void mc_replica_steps(orc_remc *s, int r, int nsteps, float *scratch_rel, float *scratch_pos,
                             double *row)
{
    const int M = s->n_mol, N = s->n_res;
    float *center = s->center + (size_t)r * M * 3;
    float *rel = s->rel + (size_t)r * N * 3;
    float *pos = s->pos + (size_t)r * N * 3;
    double *mat = s->emat + (size_t)r * M * M;
    const double rs = (double)s->rotate_step[r];
    const double rcos = cos(rs / 2.0), rsin = sin(rs / 2.0);
    for (int k = 0; k < nsteps; ++k) {
        int64_t step = s->step[r];
        int m, kind;
        double axis[3], u;
        uint32_t extra[4];
        orc_mc_proposal_flex(s->seed, r, step, M, s->flex ? s->flex->mol_flexible : NULL, &m, &kind, axis, &u,
                             extra);
        const int i0 = s->mol_off[m], i1 = s->mol_off[m + 1];
        float cnew[3] = {center[3 * m], center[3 * m + 1], center[3 * m + 2]};
        /* proposal positions live in a full-size scratch so that pair
         * indexing stays global */
        memcpy(scratch_pos, pos, sizeof(float) * 3 * N);
        memcpy(scratch_rel + 3 * i0, rel + 3 * i0, sizeof(float) * 3 * (i1 - i0));
        apply_proposal(s, r, m, kind, axis, extra, rcos, rsin, scratch_pos, scratch_rel, cnew);
        const int flexible = s->flex && s->flex->mol_flexible[m];
        double diag = 0.0;
        double delta = 0.0;
        if (flexible) {
            diag = mc_internal(s, m, scratch_pos);
            delta += diag - mat[m * M + m];
        }
        for (int J = 0; J < M; ++J) {
            if (J == m) { row[J] = 0.0; continue; }
            double lj = 0.0, dh = 0.0;
            ref_block(i0, i1, s->mol_off[J], s->mol_off[J + 1], s->type, s->charge, s->sigma, s->eps,
                      scratch_pos, s->box, &lj, &dh);
            row[J] = pair_total(lj, dh);
            delta += row[J] - mat[m * M + J];
        }
        int64_t *cnt = s->counters + 3 * r;
        const int decision = orc_metropolis(delta, s->temperature[r], u);
        cnt[decision]++;
        if (decision < 2) {
            for (int q = 0; q < 3; ++q) center[3 * m + q] = cnew[q];
            memcpy(rel + 3 * i0, scratch_rel + 3 * i0, sizeof(float) * 3 * (i1 - i0));
            memcpy(pos + 3 * i0, scratch_pos + 3 * i0, sizeof(float) * 3 * (i1 - i0));
            for (int J = 0; J < M; ++J) {
                if (J == m) continue;
                mat[m * M + J] = mat[J * M + m] = row[J];
            }
            if (flexible) mat[m * M + m] = diag;
            s->energy[r] += delta;
        }
        s->step[r] = step + 1;
    }
}

void orc_remc_mc_steps(orc_remc *s, int nsteps)
{
    const int N = s->n_res, M = s->n_mol;
    float *srel = (float *)malloc(sizeof(float) * 3 * N);
    float *spos = (float *)malloc(sizeof(float) * 3 * N);
    double *row = (double *)malloc(sizeof(double) * M);
    for (int r = 0; r < s->n_rep; ++r) mc_replica_steps(s, r, nsteps, srel, spos, row);
    free(srel); free(spos); free(row);
}

void orc_remc_exchange(orc_remc *s)
{
    const int R = s->n_rep;
    int64_t k = *s->exchange_round;
    int offset = (int)(k & 1);
    const uint32_t key[2] = {(uint32_t)s->seed, (uint32_t)(s->seed >> 32)};
    for (int ti = offset; ti < R - 1; ti += 2) {
        int i = s->rep_at_rank[ti], j = s->rep_at_rank[ti + 1];
        /* src/Simulation.cpp:780 */
        double delta = (1.0 / s->temperature[i] - 1.0 / s->temperature[j]) *
                       (s->energy[i] - s->energy[j]) * (double)(4184.0f / ORC_RGAS);
        uint32_t c[4] = {(uint32_t)k, (uint32_t)((uint64_t)k >> 32), (uint32_t)ti, ORC_TAG_EX};
        uint32_t w[4];
        orc_philox4x32_10(c, key, w);
        double u = u53(w[0], w[1]);
        double p = exp(delta);
        if (p > 1.0) p = 1.0;
        s->exchange_stats[0]++;
        if (u < p) {
            /* src/Replica.cpp:188-210: temperature, label and statistics move */
            double t = s->temperature[i]; s->temperature[i] = s->temperature[j]; s->temperature[j] = t;
            int l = s->label[i]; s->label[i] = s->label[j]; s->label[j] = l;
            for (int q = 0; q < 3; ++q) {
                int64_t a = s->counters[3 * i + q];
                s->counters[3 * i + q] = s->counters[3 * j + q];
                s->counters[3 * j + q] = a;
            }
            s->rank_of[i] = ti + 1;
            s->rank_of[j] = ti;
            s->rep_at_rank[ti] = j;
            s->rep_at_rank[ti + 1] = i;
            s->exchange_stats[1]++;
        }
    }
    *s->exchange_round = k + 1;
}

/* ------------------------------------------------------------------ */
/* Lean FP64 sampler for long CPU REMC runs                             */
/* ------------------------------------------------------------------ */
/* The Markov chain of orc_remc_mc_steps (same proposals, moves and
 * orc_metropolis decisions, delta-E rows against the molecule-pair
 * matrix) with the pair term in FP64 written for speed: a table per type
 * pair of -4E, sigma^2, the squared repulsive cutoff (or -1) and 2E,
 * (sigma/r)^6 from sigma^2 / r^2, the minimum image by rounding, and the
 * Debye-Hueckel term only over the charged residues.  Rigid molecules,
 * crowding off.  Pinned against orc_energy_fp64. */
typedef struct {
    double a[400], s2[400], rc2[400], e2[400];
    int *charged;     /* the charged residues, ascending */
    int *charged_at;  /* n_res + 1: charged residues before each residue */
} fast_tab;

static void fast_table(const orc_remc *s, fast_tab *t)
{
    const double r0 = pow(2.0, 1.0 / 6.0);
    for (int ti = 0; ti < 20; ++ti)
        for (int tj = 0; tj < 20; ++tj) {
            const int k = ti * 20 + tj;
            const double E = (double)ORC_LAMBDA * ((double)s->eps[k] - (double)ORC_E0);
            const double sij = 0.5 * ((double)s->sigma[ti] + (double)s->sigma[tj]);
            t->a[k] = -4.0 * E;
            t->s2[k] = sij * sij;
            t->rc2[k] = E > 0.0 ? (sij * r0) * (sij * r0) : -1.0;
            t->e2[k] = 2.0 * E;
        }
    t->charged = (int *)malloc(sizeof(int) * (s->n_res + 1));
    t->charged_at = (int *)malloc(sizeof(int) * (s->n_res + 1));
    int c = 0;
    for (int i = 0; i < s->n_res; ++i) {
        t->charged_at[i] = c;
        if (s->charge[s->type[i]] != 0.0f) t->charged[c++] = i;
    }
    t->charged_at[s->n_res] = c;
}

static void fast_table_free(fast_tab *t)
{
    free(t->charged);
    free(t->charged_at);
}

/* round() of the small multiples of the box that occur here, by
 * truncation (no library call) */
static inline double fast_round(double x) { return (double)(long long)(x + (x >= 0.0 ? 0.5 : -0.5)); }

/* kcal/mol of the residues [i0, i1) against [j0, j1) */
static double fast_block(const orc_remc *s, const fast_tab *t, int i0, int i1, int j0, int j1, const float *pos)
{
    const double L = (double)s->box, iL = 1.0 / L;
    double lj = 0.0, dh = 0.0;
    for (int i = i0; i < i1; ++i) {
        const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
        const int row = s->type[i] * 20;
        for (int j = j0; j < j1; ++j) {
            double dx = xi - (double)pos[3 * j], dy = yi - (double)pos[3 * j + 1], dz = zi - (double)pos[3 * j + 2];
            dx -= L * fast_round(dx * iL);
            dy -= L * fast_round(dy * iL);
            dz -= L * fast_round(dz * iL);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const int k = row + s->type[j];
            const double q = t->s2[k] / r2, s6 = q * q * q;
            const double u = t->a[k] * s6 * (s6 - 1.0);
            lj += r2 < t->rc2[k] ? t->e2[k] - u : u;
        }
    }
    /* Debye-Hueckel: charged pairs only */
    for (int a = t->charged_at[i0]; a < t->charged_at[i1]; ++a) {
        const int i = t->charged[a];
        const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
        const double qi = (double)s->charge[s->type[i]];
        for (int b = t->charged_at[j0]; b < t->charged_at[j1]; ++b) {
            const int j = t->charged[b];
            double dx = xi - (double)pos[3 * j], dy = yi - (double)pos[3 * j + 1], dz = zi - (double)pos[3 * j + 2];
            dx -= L * fast_round(dx * iL);
            dy -= L * fast_round(dy * iL);
            dz -= L * fast_round(dz * iL);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            dh += qi * (double)s->charge[s->type[j]] * exp(-r / (double)ORC_XI) / r;
        }
    }
    return pair_total(lj, dh);
}

double orc_energy_fast(const orc_remc *s, int r)
{
    fast_tab t;
    fast_table(s, &t);
    const int M = s->n_mol;
    const float *pos = s->pos + (size_t)r * s->n_res * 3;
    double e = 0.0;
    for (int I = 0; I < M; ++I)
        for (int J = I + 1; J < M; ++J)
            e += fast_block(s, &t, s->mol_off[I], s->mol_off[I + 1], s->mol_off[J], s->mol_off[J + 1], pos);
    fast_table_free(&t);
    return e;
}

void orc_remc_refresh_fast(orc_remc *s)
{
    fast_tab t;
    fast_table(s, &t);
    const int M = s->n_mol;
    for (int r = 0; r < s->n_rep; ++r) {
        const float *pos = s->pos + (size_t)r * s->n_res * 3;
        double *mat = s->emat + (size_t)r * M * M;
        double e = 0.0;
        for (int I = 0; I < M; ++I) {
            mat[I * M + I] = 0.0;
            for (int J = I + 1; J < M; ++J) {
                const double v = fast_block(s, &t, s->mol_off[I], s->mol_off[I + 1], s->mol_off[J],
                                            s->mol_off[J + 1], pos);
                mat[I * M + J] = mat[J * M + I] = v;
                e += v;
            }
        }
        s->energy[r] = e;
    }
    fast_table_free(&t);
}

typedef struct {
    orc_remc *s;
    const fast_tab *t;
    int r0, r1, nsteps;
} fast_job;

static void *fast_worker(void *arg)
{
    fast_job *job = (fast_job *)arg;
    orc_remc *s = job->s;
    const int M = s->n_mol, N = s->n_res;
    float *spos = (float *)malloc(sizeof(float) * 3 * N);
    float *srel = (float *)malloc(sizeof(float) * 3 * N);
    double *row = (double *)malloc(sizeof(double) * M);
    for (int r = job->r0; r < job->r1; ++r) {
        float *center = s->center + (size_t)r * M * 3;
        float *rel = s->rel + (size_t)r * N * 3;
        float *pos = s->pos + (size_t)r * N * 3;
        double *mat = s->emat + (size_t)r * M * M;
        const double rs = (double)s->rotate_step[r];
        const double rcos = cos(rs / 2.0), rsin = sin(rs / 2.0);
        memcpy(spos, pos, sizeof(float) * 3 * N);
        for (int k = 0; k < job->nsteps; ++k) {
            const int64_t step = s->step[r];
            int m, kind;
            double axis[3], u;
            uint32_t extra[4];
            orc_mc_proposal_flex(s->seed, r, step, M, NULL, &m, &kind, axis, &u, extra);
            const int i0 = s->mol_off[m], i1 = s->mol_off[m + 1];
            float cnew[3] = {center[3 * m], center[3 * m + 1], center[3 * m + 2]};
            memcpy(srel + 3 * i0, rel + 3 * i0, sizeof(float) * 3 * (i1 - i0));
            apply_proposal(s, r, m, kind, axis, extra, rcos, rsin, spos, srel, cnew);
            double delta = 0.0;
            for (int J = 0; J < M; ++J) {
                if (J == m) { row[J] = 0.0; continue; }
                row[J] = fast_block(s, job->t, i0, i1, s->mol_off[J], s->mol_off[J + 1], spos);
                delta += row[J] - mat[m * M + J];
            }
            const int decision = orc_metropolis(delta, s->temperature[r], u);
            s->counters[3 * r + decision]++;
            if (decision < 2) {
                for (int q = 0; q < 3; ++q) center[3 * m + q] = cnew[q];
                memcpy(rel + 3 * i0, srel + 3 * i0, sizeof(float) * 3 * (i1 - i0));
                memcpy(pos + 3 * i0, spos + 3 * i0, sizeof(float) * 3 * (i1 - i0));
                for (int J = 0; J < M; ++J)
                    if (J != m) mat[m * M + J] = mat[J * M + m] = row[J];
                s->energy[r] += delta;
            } else {
                memcpy(spos + 3 * i0, pos + 3 * i0, sizeof(float) * 3 * (i1 - i0));
            }
            s->step[r] = step + 1;
        }
    }
    free(spos); free(srel); free(row);
    return NULL;
}

void orc_remc_mc_steps_fast(orc_remc *s, int nsteps, int threads)
{
    fast_tab t;
    fast_table(s, &t);
    if (threads < 1) threads = 1;
    if (threads > s->n_rep) threads = s->n_rep;
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    fast_job *jobs = (fast_job *)malloc(sizeof(fast_job) * threads);
    int used = 0;
    for (int k = 0; k < threads; ++k) {
        const int r0 = (int)((long)k * s->n_rep / threads), r1 = (int)((long)(k + 1) * s->n_rep / threads);
        if (r0 == r1) continue;
        jobs[used].s = s; jobs[used].t = &t; jobs[used].r0 = r0; jobs[used].r1 = r1; jobs[used].nsteps = nsteps;
        pthread_create(&tid[used], NULL, fast_worker, &jobs[used]);
        ++used;
    }
    for (int k = 0; k < used; ++k) pthread_join(tid[k], NULL);
    free(tid); free(jobs);
    fast_table_free(&t);
}

/* ------------------------------------------------------------------ */
/* CPU baseline: the reference's CPU loop                               */
/* ------------------------------------------------------------------ */
typedef struct {
    orc_remc *s;
    int r0, r1, nsteps;
} ref_job;

static void *ref_worker(void *arg)
{
    ref_job *job = (ref_job *)arg;
    orc_remc *s = job->s;
    const int M = s->n_mol, N = s->n_res;
    float *brel = (float *)malloc(sizeof(float) * 3 * N);
    float *bpos = (float *)malloc(sizeof(float) * 3 * N);
    for (int r = job->r0; r < job->r1; ++r) {
        float *center = s->center + (size_t)r * M * 3;
        float *rel = s->rel + (size_t)r * N * 3;
        float *pos = s->pos + (size_t)r * N * 3;
        const double rs = (double)s->rotate_step[r];
        const double rcos = cos(rs / 2.0), rsin = sin(rs / 2.0);
        for (int k = 0; k < job->nsteps; ++k) {
            int64_t step = s->step[r];
            int m, kind;
            double axis[3], u;
            uint32_t extra[4];
            orc_mc_proposal_flex(s->seed, r, step, M, s->flex ? s->flex->mol_flexible : NULL, &m, &kind, axis,
                                 &u, extra);
            const int i0 = s->mol_off[m], i1 = s->mol_off[m + 1];
            /* src/Replica.cpp:357 MC_backup_restore */
            float bc[3] = {center[3 * m], center[3 * m + 1], center[3 * m + 2]};
            memcpy(brel + 3 * i0, rel + 3 * i0, sizeof(float) * 3 * (i1 - i0));
            memcpy(bpos + 3 * i0, pos + 3 * i0, sizeof(float) * 3 * (i1 - i0));
            apply_proposal(s, r, m, kind, axis, extra, rcos, rsin, pos, rel, center + 3 * m);
            /* src/Replica.cpp:373-380 full recompute, E() + internal_molecule_E(true) */
            double lj, dh;
            orc_energy_ref(M, s->mol_off, s->type, s->charge, s->sigma, s->eps, pos, s->box, &lj, &dh);
            double enew = pair_total(lj, dh);
            if (s->flex)
                for (int I = 0; I < M; ++I)
                    if (s->flex->mol_flexible[I]) {
                        double v[5];
                        orc_internal_ref(s->flex, I, s->mol_off, s->type, s->charge, s->sigma, s->eps, pos,
                                         s->box, 0, v);
                        enew += orc_internal_total(v);
                    }
            /* src/Replica.cpp:385-410 with float delta and float exponent */
            const int decision = orc_metropolis_ref((float)(enew - s->energy[r]), (float)s->temperature[r], u);
            int64_t *cnt = s->counters + 3 * r;
            cnt[decision]++;
            if (decision < 2) {
                s->energy[r] = enew;
            } else {
                for (int q = 0; q < 3; ++q) center[3 * m + q] = bc[q];
                memcpy(rel + 3 * i0, brel + 3 * i0, sizeof(float) * 3 * (i1 - i0));
                memcpy(pos + 3 * i0, bpos + 3 * i0, sizeof(float) * 3 * (i1 - i0));
            }
            s->step[r] = step + 1;
        }
    }
    free(brel); free(bpos);
    return NULL;
}

double orc_remc_reference_run(orc_remc *s, int nsteps, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > s->n_rep) threads = s->n_rep;
    pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    ref_job *jobs = (ref_job *)malloc(sizeof(ref_job) * threads);
    /* src/Simulation.cpp:479-487,684-688: contiguous blocks of replicas */
    int per = (s->n_rep + threads - 1) / threads;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int used = 0;
    for (int t = 0; t < threads; ++t) {
        int r0 = t * per, r1 = r0 + per;
        if (r0 >= s->n_rep) break;
        if (r1 > s->n_rep) r1 = s->n_rep;
        jobs[t].s = s; jobs[t].r0 = r0; jobs[t].r1 = r1; jobs[t].nsteps = nsteps;
        pthread_create(&tid[t], NULL, ref_worker, &jobs[t]);
        used++;
    }
    for (int t = 0; t < used; ++t) pthread_join(tid[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(tid); free(jobs);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
