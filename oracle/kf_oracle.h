/*
 * kf_oracle.h -- CPU oracle for the Kern-Frenkel rototranslation Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a strict-IEEE C restatement of the
 * reference algorithm (PatchyParticles, C99).  Only tests/, the smoke check
 * and the cpu_baseline leg of bench.py may load it; the product library
 * (libpp_hip.so) never links or calls it.
 *
 * Build (oracle/Makefile): gcc -O2 -ffp-contract=off, no fast-math, so that
 * every multiply and add rounds exactly where the reference source writes it.
 *
 * Every function cites the reference file:line it restates.  Pinning status
 * of each function (what checks it beyond retyping the formula) is listed in
 * the header comment of kf_oracle.c.
 */
#ifndef KF_ORACLE_H
#define KF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Interaction results, as in defs.h:31-33. */
#define ORC_OVERLAP (-1)
#define ORC_NO_BOND 0
#define ORC_BOND 1

/* ---- geometry (snapshot level) ---------------------------------------- */

/* 4 tetrahedral base patches, each normalised (system.c:20-31, utils.c:31-36). */
void orc_base_patches(double base[12]);

/* patch_i = O * base_i, rows of O, left-to-right sums (utils.h:17-21). */
void orc_patches(const double O[9], const double base[12], double P[12]);

/* Kern-Frenkel pair test between p (rp, Pp) and q (rq, Pq) (MC.c:232-266).
 * Returns ORC_OVERLAP / ORC_NO_BOND / ORC_BOND; onp and onq receive the bonding
 * patch pair (may be NULL). */
int orc_pair(const double box[3], double rc2, double cosmax,
             const double rp[3], const double Pp[12],
             const double rq[3], const double Pq[12], int *onp, int *onq);

/* Gram-Schmidt on the three rows of O, then normalise each (utils.c:91-113). */
void orc_gram_schmidt(double O[9]);

/* Reference cell index of a position (cells.c:43-49). nside[3] per axis. */
int orc_cell_index(const double r[3], const double box[3], const int nside[3], int idx[3]);

/* Reference grid: N_side = max(3, floor(L / rcut)) (cells.c:21-28). */
void orc_grid(const double box[3], double rcut, int nside[3]);

/* Per-particle MC_energy semantics on a snapshot, with the reference 27-cell
 * linked-cell scan (MC.c:272-309): bonds[i] = number of bonded neighbours of
 * i; overlap[i] = 1 if i overlaps any neighbour (then bonds[i] = 0, as the
 * reference returns 0 on the first overlap).  Returns the total number of
 * bonded pairs (sum/2) or -1 if any overlap. */
long long orc_energy_cells(int n, const double *r, const double *O, const double box[3],
                           double delta, double cosmax, int32_t *bonds, uint8_t *overlap);

/* Same quantity by O(N^2) brute force over all pairs (no cells). */
long long orc_energy_brute(int n, const double *r, const double *O, const double box[3],
                           double delta, double cosmax, int32_t *bonds, uint8_t *overlap);

/* ---- the serial reference-algorithm sampler -------------------------- */

typedef struct orc_params {
	int ensemble;              /* 0 NVT, 5 NPT (defs.h:21-25) */
	double temperature;        /* Temperature */
	double disp_max;           /* Disp_max */
	double theta_max;          /* Theta_max */
	double kf_delta;           /* KF_delta */
	double kf_cosmax;          /* KF_cosmax */
	double pressure;           /* P (NPT) */
	double rescale_factor_max; /* rescale_factor_max (NPT) */
	long seed;                 /* Seed, srand48 semantics (system.c:71-75) */
} orc_params;

typedef struct orc_sim orc_sim;

/* Build a system from in-memory positions (3n) and orientation matrices (9n,
 * rows), as system_init does after reading a configuration: patches are
 * O*base, cells are built with r_cut = 1 + delta (system.c:198-251). */
orc_sim *orc_sim_create(const orc_params *prm, int n, const double *r, const double *O,
                        const double box[3]);
void orc_sim_destroy(orc_sim *s);

/* Initial energy as main.c:54-61; returns 1 on overlap. */
int orc_sim_init_energy(orc_sim *s);

/* n calls of do_NVT / do_NPT (MC.c:19-29, 60-68). */
void orc_sim_sweeps(orc_sim *s, long long n);

/* MC_check_energy (MC.c:766-784): returns 0 ok, 1 overlap, 2 drift. */
int orc_sim_check_energy(orc_sim *s);

/* cells_check (cells.c:65-76): number of particles reachable from heads. */
int orc_sim_cells_count(orc_sim *s);

double orc_sim_energy(const orc_sim *s);
void orc_sim_get(const orc_sim *s, double *r, double *O, double box[3]);
/* tries/accepted for [ROTO_TRASL, VOLUME] since the last reset (defs.h:11-19). */
void orc_sim_counters(const orc_sim *s, long long tries[2], long long accepted[2]);
void orc_sim_reset_counters(orc_sim *s);

/* ---- initial configurations ------------------------------------------- */

/* Random sequential insertion of n spheres at density rho (generator.c:10-124).
 * seed < 0 keeps the default drand48 state (the reference generator never
 * seeds).  Writes r (3n), O (9n), box (3).  Returns 0, or -1 if rho > 0.7. */
int orc_generate(int n, double rho, long seed, double *r, double *O, double box[3]);

/* Load a .rrr configuration (system.c:77-81, 213-245): header
 * "step N Lx Ly Lz"; 3 lines per particle (O row 0, O row 1, position);
 * rows 0 and 1 normalised, row 2 = row0 x row1, then Gram-Schmidt.
 * Pass r == NULL to read only the header. Returns N or -1. */
int orc_rrr_load(const char *path, long long *step, double box[3], double *r, double *O, int n_cap);

/* Write a .rrr (output.c:302-317): header %lld %d %.12lf x3; rows with %lf,
 * positions with %.12lf. */
int orc_rrr_save(const char *path, long long step, int n, const double box[3],
                 const double *r, const double *O);

/* generated.rrr writer: as orc_rrr_save but with the generator header
 * "0 N Lx Ly Lz" with %lf box (generator.c:73). */
int orc_rrr_save_generated(const char *path, int n, const double box[3], const double *r, const double *O);

/* Philox4x32-10 (Salmon et al., SC'11), the counter RNG of the GPU path:
 * out = bijection(ctr; key). Restated here as the checker for the device. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* n blocks at once: ctr[4n], key[2n] -> out[4n]. */
void orc_philox4x32_10_n(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
