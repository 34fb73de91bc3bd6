/*
 * kf_oracle.c -- CPU oracle: strict-IEEE restatement of the reference's
 * Kern-Frenkel rototranslation Monte Carlo (PatchyParticles).
 *
 * TEST INFRASTRUCTURE ONLY (tests/, smoke check, bench cpu_baseline leg).
 * The product path (libpp_hip.so) never links or calls this file.
 *
 * Compile with -ffp-contract=off and without -ffast-math: every expression
 * below is written in the reference's operation order so that the doubles it
 * produces are the bits the reference produces when it is compiled
 * strict-IEEE.  drand48 is used through erand48() on a private state, which is
 * the same 48-bit LCG and the same 48-bit -> double conversion as the global
 * drand48() the reference calls.
 *
 * Pinning (what checks each part, beyond retyping the formula):
 *  - orc_base_patches / orc_patches / orc_pair / orc_energy_*:
 *      per-particle bond counts and overlap flags equal to the reference's own
 *      MC_energy (built from its sources into oracle/_ref by oracle/Makefile)
 *      on committed snapshots (the .npz files of tests/golden); hand-built known answers
 *      at the strict thresholds (d^2 = 1, d^2 = (1+delta)^2, cos = cosmax);
 *      pair symmetry; brute force O(N^2) == cell list.
 *  - orc_sim_* (the sampler): bit-exact trajectories (positions and
 *      orientation matrices after k sweeps, NVT and NPT) against the
 *      reference executable's code run through oracle/_ref (golden fixtures);
 *      two-particle bonding probability closed form.
 *  - orc_generate: byte-identical generated.rrr against the reference
 *      generator for the same N and rho.
 *  - orc_rrr_load / orc_rrr_save: round trip against the reference loader
 *      (energies and states after loading the same file).
 *  - orc_philox4x32_10: Random123 known-answer vectors and rocRAND.
 */
#include "kf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* small vector helpers, operation order as utils.h:14-15 / utils.c:24-36  */

static inline double dot3(const double *a, const double *b) {
	/* SCALAR(a, b) = a0*b0 + a1*b1 + a2*b2, summed left to right */
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline void unit3(double *v) {
	/* normalize(): one sqrt, then three divisions (utils.c:31-36) */
	double len = sqrt(dot3(v, v));
	v[0] /= len;
	v[1] /= len;
	v[2] /= len;
}

static inline void cross3(const double *a, const double *b, double *c) {
	/* cross() (utils.c:24-28) */
	c[0] = a[1] * b[2] - a[2] * b[1];
	c[1] = a[2] * b[0] - a[0] * b[2];
	c[2] = a[0] * b[1] - a[1] * b[0];
}

void orc_base_patches(double base[12]) {
	/* system.c:20-31: (-h,-h,h), (h,-h,-h), (h,h,h), (-h,h,-h), h = 0.5/sqrt(3) */
	const double h = 0.5 / sqrt(3.);
	const double sgn[4][3] = {{-1, -1, 1}, {1, -1, -1}, {1, 1, 1}, {-1, 1, -1}};
	for(int i = 0; i < 4; i++) {
		for(int k = 0; k < 3; k++) base[3 * i + k] = sgn[i][k] > 0 ? h : -h;
		unit3(base + 3 * i);
	}
}

void orc_patches(const double O[9], const double base[12], double P[12]) {
	/* MATRIX_VECTOR_MULTIPLICATION(O, b, p) (utils.h:17-21), O stored by rows */
	for(int i = 0; i < 4; i++) {
		const double *b = base + 3 * i;
		for(int k = 0; k < 3; k++) P[3 * i + k] = O[3 * k + 0] * b[0] + O[3 * k + 1] * b[1] + O[3 * k + 2] * b[2];
	}
}

int orc_pair(const double box[3], double rc2, double cosmax, const double rp[3], const double Pp[12],
             const double rq[3], const double Pq[12], int *onp, int *onq) {
	/* MC_would_interact (MC.c:232-266): d = r_q - r_p, minimum image by rint */
	double d[3];
	for(int k = 0; k < 3; k++) {
		d[k] = rq[k] - rp[k];
		d[k] -= box[k] * rint(d[k] / box[k]);
	}
	double d2 = dot3(d, d);
	if(d2 < 1.) return ORC_OVERLAP;
	if(d2 < rc2) {
		double len = sqrt(d2);
		d[0] /= len;
		d[1] /= len;
		d[2] /= len;
		for(int a = 0; a < 4; a++) {
			double cp = dot3(d, Pp + 3 * a);
			if(cp > cosmax) {
				for(int b = 0; b < 4; b++) {
					double cq = -dot3(d, Pq + 3 * b);
					if(cq > cosmax) {
						if(onp) *onp = a;
						if(onq) *onq = b;
						return ORC_BOND;
					}
				}
			}
		}
	}
	return ORC_NO_BOND;
}

void orc_gram_schmidt(double O[9]) {
	/* gram_schmidt(v1, v2, v3) on the rows of O (utils.c:91-113) */
	double *v1 = O, *v2 = O + 3, *v3 = O + 6;
	double n1 = dot3(v1, v1);
	double c21 = dot3(v2, v1) / n1;
	v2[0] -= c21 * v1[0];
	v2[1] -= c21 * v1[1];
	v2[2] -= c21 * v1[2];
	double s31 = dot3(v3, v1);
	double s32 = dot3(v3, v2);
	double n2 = dot3(v2, v2);
	double c31 = s31 / n1;
	double c32 = s32 / n2;
	v3[0] -= c31 * v1[0] + c32 * v2[0];
	v3[1] -= c31 * v1[1] + c32 * v2[1];
	v3[2] -= c31 * v1[2] + c32 * v2[2];
	unit3(v1);
	unit3(v2);
	unit3(v3);
}

int orc_cell_index(const double r[3], const double box[3], const int nside[3], int idx[3]) {
	/* cells_fill_and_get_idx_from_vector (cells.c:43-49) */
	for(int k = 0; k < 3; k++) {
		double t = r[k] / box[k];
		idx[k] = (int) ((t - floor(t)) * (1. - DBL_EPSILON) * nside[k]);
	}
	return (idx[0] * nside[1] + idx[1]) * nside[2] + idx[2];
}

void orc_grid(const double box[3], double rcut, int nside[3]) {
	/* cells_init (cells.c:21-28) */
	for(int k = 0; k < 3; k++) {
		nside[k] = floor(box[k] / rcut);
		if(nside[k] < 3) nside[k] = 3;
	}
}

/* ---------------------------------------------------------------------- */
/* linked cells (cells.c), particle indices instead of pointers             */

typedef struct {
	int n[3], ncell;
	int *head; /* ncell, -1 = empty */
	int *next; /* per particle, -1 = end */
} lcells;

static void lc_init(lcells *c, const double box[3], double rcut, int nmax) {
	orc_grid(box, rcut, c->n);
	c->ncell = c->n[0] * c->n[1] * c->n[2];
	c->head = malloc(sizeof(int) * c->ncell);
	c->next = malloc(sizeof(int) * (nmax > 0 ? nmax : 1));
	for(int i = 0; i < c->ncell; i++) c->head[i] = -1;
	for(int i = 0; i < nmax; i++) c->next[i] = -1;
}

static void lc_free(lcells *c) {
	free(c->head);
	free(c->next);
	c->head = c->next = NULL;
}

/* ---------------------------------------------------------------------- */
/* snapshot energies                                                       */

static int scan_particle(const lcells *c, int i, const double *r, const double *P, const double box[3],
                         double rc2, double cosmax, int *bonds) {
	/* MC_energy (MC.c:272-309): 27 cells around i, x outer, z inner */
	int ind[3];
	orc_cell_index(r + 3 * i, box, c->n, ind);
	int nb = 0;
	for(int j = -1; j < 2; j++) {
		int cx = (ind[0] + j + c->n[0]) % c->n[0];
		for(int k = -1; k < 2; k++) {
			int cy = (ind[1] + k + c->n[1]) % c->n[1];
			for(int l = -1; l < 2; l++) {
				int cz = (ind[2] + l + c->n[2]) % c->n[2];
				int q = c->head[(cx * c->n[1] + cy) * c->n[2] + cz];
				while(q != -1) {
					if(q != i) {
						int v = orc_pair(box, rc2, cosmax, r + 3 * i, P + 12 * i, r + 3 * q, P + 12 * q, NULL, NULL);
						if(v == ORC_BOND) nb++;
						else if(v == ORC_OVERLAP) {
							*bonds = 0;
							return 1;
						}
					}
					q = c->next[q];
				}
			}
		}
	}
	*bonds = nb;
	return 0;
}

static double *make_patches(int n, const double *O) {
	double base[12];
	orc_base_patches(base);
	double *P = malloc(sizeof(double) * 12 * (n > 0 ? n : 1));
	for(int i = 0; i < n; i++) orc_patches(O + 9 * i, base, P + 12 * i);
	return P;
}

long long orc_energy_cells(int n, const double *r, const double *O, const double box[3], double delta,
                           double cosmax, int32_t *bonds, uint8_t *overlap) {
	double rcut = 1. + delta;
	double rc2 = rcut * rcut; /* SQR(1. + delta) (system.c:86) */
	double *P = make_patches(n, O);
	lcells c;
	lc_init(&c, box, rcut, n);
	int ind[3];
	for(int i = 0; i < n; i++) { /* cells_fill (cells.c:51-63): push front */
		int ci = orc_cell_index(r + 3 * i, box, c.n, ind);
		c.next[i] = c.head[ci];
		c.head[ci] = i;
	}
	long long tot = 0;
	int any = 0;
	for(int i = 0; i < n; i++) {
		int b;
		int ov = scan_particle(&c, i, r, P, box, rc2, cosmax, &b);
		if(bonds) bonds[i] = b;
		if(overlap) overlap[i] = (uint8_t) ov;
		any |= ov;
		tot += b;
	}
	lc_free(&c);
	free(P);
	return any ? -1 : tot / 2;
}

long long orc_energy_brute(int n, const double *r, const double *O, const double box[3], double delta,
                           double cosmax, int32_t *bonds, uint8_t *overlap) {
	double rcut = 1. + delta;
	double rc2 = rcut * rcut;
	double *P = make_patches(n, O);
	long long tot = 0;
	int any = 0;
	for(int i = 0; i < n; i++) {
		int nb = 0, ov = 0;
		for(int q = 0; q < n; q++) {
			if(q == i) continue;
			int v = orc_pair(box, rc2, cosmax, r + 3 * i, P + 12 * i, r + 3 * q, P + 12 * q, NULL, NULL);
			if(v == ORC_OVERLAP) ov = 1;
			else if(v == ORC_BOND) nb++;
		}
		if(ov) nb = 0;
		if(bonds) bonds[i] = nb;
		if(overlap) overlap[i] = (uint8_t) ov;
		any |= ov;
		tot += nb;
	}
	free(P);
	return any ? -1 : tot / 2;
}

/* ---------------------------------------------------------------------- */
/* the serial sampler                                                      */

enum { MV_ROTO = 0, MV_VOLUME = 6, N_MV = 8 }; /* defs.h:11-19 */

struct orc_sim {
	orc_params prm;
	int n;
	double box[3], V;
	double base[12], rcut, rc2;
	double *r, *r_old, *O, *O_old, *P;
	int *cell, *cell_old;
	lcells cells;
	unsigned short xs[3]; /* drand48 state */
	double energy;
	int overlap;
	long long tries[N_MV], accepted[N_MV];
};

static inline double rnd(orc_sim *s) { return erand48(s->xs); }

static void sphere_point(orc_sim *s, double v[3]);

static void lcg48_seed(unsigned short xs[3], long seed) {
	/* srand48(seed): X = (seed & 0xffffffff) << 16 | 0x330e */
	unsigned long v = (unsigned long) seed & 0xffffffffUL;
	xs[0] = 0x330e;
	xs[1] = (unsigned short) (v & 0xffff);
	xs[2] = (unsigned short) (v >> 16);
}

static void sim_fill_cells(orc_sim *s) {
	/* cells_init + cells_fill (cells.c:17-63) */
	lc_init(&s->cells, s->box, s->rcut, s->n);
	int ind[3];
	for(int i = 0; i < s->n; i++) {
		int ci = orc_cell_index(s->r + 3 * i, s->box, s->cells.n, ind);
		s->cells.next[i] = s->cells.head[ci];
		s->cells.head[ci] = i;
		s->cell[i] = ci;
		s->cell_old[i] = ci;
	}
}

orc_sim *orc_sim_create(const orc_params *prm, int n, const double *r, const double *O, const double box[3]) {
	orc_sim *s = calloc(1, sizeof(orc_sim));
	s->prm = *prm;
	s->n = n;
	memcpy(s->box, box, sizeof(double) * 3);
	s->V = box[0] * box[1] * box[2];
	orc_base_patches(s->base);
	s->rc2 = (1. + prm->kf_delta) * (1. + prm->kf_delta);
	s->rcut = 1. + prm->kf_delta;
	size_t m = n > 0 ? n : 1;
	s->r = malloc(sizeof(double) * 3 * m);
	s->r_old = malloc(sizeof(double) * 3 * m);
	s->O = malloc(sizeof(double) * 9 * m);
	s->O_old = malloc(sizeof(double) * 9 * m);
	s->P = malloc(sizeof(double) * 12 * m);
	s->cell = malloc(sizeof(int) * m);
	s->cell_old = malloc(sizeof(int) * m);
	memcpy(s->r, r, sizeof(double) * 3 * n);
	memcpy(s->O, O, sizeof(double) * 9 * n);
	for(int i = 0; i < n; i++) orc_patches(s->O + 9 * i, s->base, s->P + 12 * i);
	lcg48_seed(s->xs, prm->seed);
	/* system_init seeds drand48 (system.c:75) before _init_tetrahedral_patches,
	 * whose set_orientation_around_vector call draws two sphere points
	 * (system.c:48 -> utils.c:224-225); the values are unused on this path,
	 * the draws still advance the stream. */
	double unused[3];
	sphere_point(s, unused);
	sphere_point(s, unused);
	sim_fill_cells(s);
	return s;
}

void orc_sim_destroy(orc_sim *s) {
	if(!s) return;
	lc_free(&s->cells);
	free(s->r);
	free(s->r_old);
	free(s->O);
	free(s->O_old);
	free(s->P);
	free(s->cell);
	free(s->cell_old);
	free(s);
}

static double sim_particle_energy(orc_sim *s, int i) {
	/* MC_energy (MC.c:272-309), with the overlap side channel */
	s->overlap = 0;
	int b;
	if(scan_particle(&s->cells, i, s->r, s->P, s->box, s->rc2, s->prm.kf_cosmax, &b)) {
		s->overlap = 1;
		return 0.;
	}
	return -(double) b;
}

int orc_sim_init_energy(orc_sim *s) {
	/* main.c:54-61 */
	s->energy = 0;
	for(int i = 0; i < s->n; i++) {
		s->energy += sim_particle_energy(s, i);
		if(s->overlap == 1) return 1;
	}
	s->energy *= 0.5;
	return 0;
}

static void unlink_from(orc_sim *s, int c, int i) {
	int *head = s->cells.head, *next = s->cells.next;
	if(head[c] == i) {
		head[c] = next[i];
		return;
	}
	int q = head[c];
	while(next[q] != i) q = next[q];
	next[q] = next[i];
}

static void change_cell(orc_sim *s, int i) {
	/* MC_change_cell (MC.c:180-206) */
	int ind[3];
	int c = orc_cell_index(s->r + 3 * i, s->box, s->cells.n, ind);
	if(c == s->cell[i]) {
		s->cell_old[i] = s->cell[i];
		return;
	}
	unlink_from(s, s->cell[i], i);
	s->cells.next[i] = s->cells.head[c];
	s->cells.head[c] = i;
	s->cell_old[i] = s->cell[i];
	s->cell[i] = c;
}

static void rollback(orc_sim *s, int i) {
	/* MC_rollback_particle (MC.c:146-178) */
	memcpy(s->r + 3 * i, s->r_old + 3 * i, sizeof(double) * 3);
	memcpy(s->O + 9 * i, s->O_old + 9 * i, sizeof(double) * 9);
	orc_patches(s->O + 9 * i, s->base, s->P + 12 * i);
	if(s->cell[i] != s->cell_old[i]) {
		unlink_from(s, s->cell[i], i);
		int c_old = s->cell_old[i];
		s->cells.next[i] = s->cells.head[c_old];
		s->cells.head[c_old] = i;
		s->cell_old[i] = s->cell[i];
		s->cell[i] = c_old;
	}
}

static void sphere_point(orc_sim *s, double v[3]) {
	/* random_vector_on_sphere (utils.c:115-131), Marsaglia's rejection */
	double u1, u2, q;
	do {
		u1 = 1. - 2. * rnd(s);
		u2 = 1. - 2. * rnd(s);
		q = u1 * u1 + u2 * u2;
	} while(q >= 1.);
	double f = 2. * sqrt(1. - q);
	v[0] = u1 * f;
	v[1] = u2 * f;
	v[2] = 1. - 2. * q;
}

static void rotation(const double a[3], double t, double R[9]) {
	/* get_rotation_matrix (utils.c:186-202): Rodrigues' formula */
	double st = sin(t);
	double ct = cos(t);
	double uct = 1. - ct;
	R[0] = a[0] * a[0] + (1. - a[0] * a[0]) * ct;
	R[1] = a[0] * a[1] * uct - a[2] * st;
	R[2] = a[0] * a[2] * uct + a[1] * st;
	R[3] = a[0] * a[1] * uct + a[2] * st;
	R[4] = a[1] * a[1] + (1. - a[1] * a[1]) * ct;
	R[5] = a[1] * a[2] * uct - a[0] * st;
	R[6] = a[0] * a[2] * uct - a[1] * st;
	R[7] = a[1] * a[2] * uct + a[0] * st;
	R[8] = a[2] * a[2] + (1. - a[2] * a[2]) * ct;
}

static void matmul3(const double A[9], const double B[9], double C[9]) {
	/* matrix_matrix_multiplication (utils.c:172-184) */
	for(int i = 0; i < 3; i++)
		for(int j = 0; j < 3; j++) C[3 * i + j] = A[3 * i + 0] * B[0 + j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

static void move_rototranslate(orc_sim *s) {
	/* MC_move_rototranslate (MC.c:311-340) */
	int i = (int) (rnd(s) * s->n);
	s->tries[MV_ROTO]++;
	double disp[3];
	disp[0] = (rnd(s) - 0.5) * s->prm.disp_max;
	disp[1] = (rnd(s) - 0.5) * s->prm.disp_max;
	disp[2] = (rnd(s) - 0.5) * s->prm.disp_max;
	double theta = rnd(s) * s->prm.theta_max;
	double axis[3], R[9], Onew[9];
	sphere_point(s, axis);
	rotation(axis, theta, R);
	matmul3(s->O + 9 * i, R, Onew); /* utils_rotate_matrix: O' = O R (utils.c:204-209) */

	double dE = -sim_particle_energy(s, i);
	/* MC_rototraslate_particle (MC.c:208-230) */
	double *r = s->r + 3 * i, *ro = s->r_old + 3 * i;
	ro[0] = r[0];
	ro[1] = r[1];
	ro[2] = r[2];
	r[0] += disp[0];
	r[1] += disp[1];
	r[2] += disp[2];
	memcpy(s->O_old + 9 * i, s->O + 9 * i, sizeof(double) * 9);
	memcpy(s->O + 9 * i, Onew, sizeof(double) * 9);
	orc_patches(s->O + 9 * i, s->base, s->P + 12 * i);
	change_cell(s, i);
	dE += sim_particle_energy(s, i);

	if(!s->overlap && (dE < 0. || rnd(s) < exp(-dE / s->prm.temperature))) {
		s->energy += dE;
		s->accepted[MV_ROTO]++;
	}
	else {
		rollback(s, i);
		s->overlap = 0;
	}
}

static void change_volume(orc_sim *s) {
	/* MC_change_volume (MC.c:450-513): one random axis is rescaled */
	s->tries[MV_VOLUME]++;
	double dE = -s->energy;
	double V0 = s->box[0] * s->box[1] * s->box[2];
	int dir = rnd(s) * 3;
	double lnV1 = log(V0) + (rnd(s) - 0.5) * s->prm.rescale_factor_max;
	double V1 = exp(lnV1);
	double old_side = s->box[dir];
	double new_side = V1;
	for(int k = 0; k < 3; k++)
		if(k != dir) new_side /= s->box[k];
	s->box[dir] = new_side;
	double f = new_side / old_side;
	for(int i = 0; i < s->n; i++) s->r[3 * i + dir] *= f;

	if((s->box[dir] / s->cells.n[dir]) < s->rcut) {
		lc_free(&s->cells);
		sim_fill_cells(s);
	}

	int ov = 0;
	for(int i = 0; i < s->n && !ov; i++) {
		dE += sim_particle_energy(s, i) * 0.5;
		if(s->overlap) ov = 1;
	}
	double dV = V1 - V0;
	double arg = -(s->prm.pressure * dV + dE) / s->prm.temperature + (s->n + 1) * log(f);
	if(!ov && rnd(s) < exp(arg)) {
		s->V = V1;
		s->energy += dE;
		s->accepted[MV_VOLUME]++;
	}
	else {
		for(int i = 0; i < s->n; i++) s->r[3 * i + dir] /= f;
		s->box[dir] = old_side;
		s->overlap = 0;
	}
}

void orc_sim_sweeps(orc_sim *s, long long nsweeps) {
	for(long long t = 0; t < nsweeps; t++) {
		if(s->prm.ensemble == 5) {
			/* do_NPT (MC.c:60-68) */
			for(int i = 0; i < s->n; i++) {
				if(rnd(s) < 1. / s->n) change_volume(s);
				else if(s->n > 0) move_rototranslate(s);
			}
		}
		else {
			/* do_NVT (MC.c:19-29): one extra draw per move for the Lx test */
			for(int i = 0; i < s->n; i++) {
				move_rototranslate(s);
				(void) rnd(s);
			}
		}
	}
}

int orc_sim_check_energy(orc_sim *s) {
	/* MC_check_energy (MC.c:766-784).  Patch vectors are left as they are:
	 * only the orientation rows are re-orthonormalised. */
	double E = 0.;
	int any = 0;
	for(int i = 0; i < s->n; i++) {
		E += sim_particle_energy(s, i);
		any |= s->overlap;
		orc_gram_schmidt(s->O + 9 * i);
	}
	E *= 0.5;
	if(any) return 1;
	if(fabs(s->energy) > 1e-5 && fabs((E - s->energy) / s->energy) > 1e-5) return 2;
	s->energy = E;
	return 0;
}

int orc_sim_cells_count(orc_sim *s) {
	int cnt = 0;
	for(int c = 0; c < s->cells.ncell; c++)
		for(int q = s->cells.head[c]; q != -1; q = s->cells.next[q]) cnt++;
	return cnt;
}

double orc_sim_energy(const orc_sim *s) { return s->energy; }

void orc_sim_get(const orc_sim *s, double *r, double *O, double box[3]) {
	if(r) memcpy(r, s->r, sizeof(double) * 3 * s->n);
	if(O) memcpy(O, s->O, sizeof(double) * 9 * s->n);
	if(box) memcpy(box, s->box, sizeof(double) * 3);
}

void orc_sim_counters(const orc_sim *s, long long tries[2], long long accepted[2]) {
	tries[0] = s->tries[MV_ROTO];
	tries[1] = s->tries[MV_VOLUME];
	accepted[0] = s->accepted[MV_ROTO];
	accepted[1] = s->accepted[MV_VOLUME];
}

void orc_sim_reset_counters(orc_sim *s) {
	memset(s->tries, 0, sizeof(s->tries));
	memset(s->accepted, 0, sizeof(s->accepted));
}

/* ---------------------------------------------------------------------- */
/* initial configuration (generator.c)                                     */

static double det3(const double m[9]) {
	/* determinant (utils.c:78-89), accumulated in the same order */
	double det = 0.;
	det += m[0] * m[4] * m[8];
	det += m[1] * m[5] * m[6];
	det += m[2] * m[3] * m[7];
	det -= m[2] * m[4] * m[6];
	det -= m[0] * m[5] * m[7];
	det -= m[1] * m[3] * m[8];
	return det;
}

static void random_orientation(orc_sim *s, double O[9]) {
	/* random_orientation (utils.c:133-169): rows (z, x, y), or (x, z, y) if
	 * that has a negative determinant */
	double x[3], y[3], z[3], M[9];
	sphere_point(s, x);
	sphere_point(s, y);
	sphere_point(s, z);
	memcpy(M, x, sizeof x);
	memcpy(M + 3, y, sizeof y);
	memcpy(M + 6, z, sizeof z);
	orc_gram_schmidt(M);
	memcpy(x, M, sizeof x);
	memcpy(y, M + 3, sizeof y);
	memcpy(z, M + 6, sizeof z);
	memcpy(O, z, sizeof z);
	memcpy(O + 3, x, sizeof x);
	memcpy(O + 6, y, sizeof y);
	if(det3(O) < 0) {
		memcpy(O, x, sizeof x);
		memcpy(O + 3, z, sizeof z);
	}
}

int orc_generate(int n, double rho, long seed, double *r, double *O, double box[3]) {
	if(rho > 0.7) return -1;
	orc_sim s;
	memset(&s, 0, sizeof s);
	if(seed >= 0) lcg48_seed(s.xs, seed);
	double L = pow(n / rho, 1. / 3.); /* generator.c:103-105 */
	box[0] = box[1] = box[2] = L;
	lcells c;
	lc_init(&c, box, 1., n); /* cells_init(rcut = 1) (generator.c:109) */
	int ind[3];
	int inserted = 0;
	while(inserted < n) {
		/* make_initial_conf / would_overlap (generator.c:10-68) */
		double t[3];
		t[0] = rnd(&s) * box[0];
		t[1] = rnd(&s) * box[1];
		t[2] = rnd(&s) * box[2];
		double pt[3] = {0. + t[0], 0. + t[1], 0. + t[2]};
		orc_cell_index(pt, box, c.n, ind);
		int hit = 0;
		for(int j = -1; j < 2 && !hit; j++) {
			int cx = (ind[0] + j + c.n[0]) % c.n[0];
			for(int k = -1; k < 2 && !hit; k++) {
				int cy = (ind[1] + k + c.n[1]) % c.n[1];
				for(int l = -1; l < 2 && !hit; l++) {
					int cz = (ind[2] + l + c.n[2]) % c.n[2];
					for(int q = c.head[(cx * c.n[1] + cy) * c.n[2] + cz]; q != -1; q = c.next[q]) {
						if(q == inserted) continue;
						double d[3];
						for(int a = 0; a < 3; a++) {
							d[a] = r[3 * q + a] - pt[a];
							d[a] -= box[a] * rint(d[a] / box[a]);
						}
						if(dot3(d, d) < 1.) {
							hit = 1;
							break;
						}
					}
				}
			}
		}
		if(!hit) {
			random_orientation(&s, O + 9 * inserted);
			memcpy(r + 3 * inserted, t, sizeof t);
			int ci = orc_cell_index(r + 3 * inserted, box, c.n, ind);
			c.next[inserted] = c.head[ci];
			c.head[ci] = inserted;
			inserted++;
		}
	}
	lc_free(&c);
	return 0;
}

/* ---------------------------------------------------------------------- */
/* .rrr files                                                              */

int orc_rrr_load(const char *path, long long *step, double box[3], double *r, double *O, int n_cap) {
	FILE *f = fopen(path, "r");
	if(!f) return -1;
	long long st = 0;
	int n;
	if(fscanf(f, "%lld %d %lf %lf %lf\n", &st, &n, box, box + 1, box + 2) != 5) {
		fclose(f);
		return -1;
	}
	if(step) *step = st;
	if(r == NULL) {
		fclose(f);
		return n;
	}
	char line[512];
	int i = 0;
	/* system.c:213-245: fgets row 0, fscanf row 1 and position */
	while(fgets(line, sizeof line, f) != NULL) {
		if(i >= n_cap) {
			fclose(f);
			return -1;
		}
		double a[3], b[3], c[3], p[3];
		if(sscanf(line, "%lf %lf %lf", a, a + 1, a + 2) != 3) break;
		if(fscanf(f, "%lf %lf %lf\n", b, b + 1, b + 2) != 3) break;
		if(fscanf(f, "%lf %lf %lf\n", p, p + 1, p + 2) != 3) break;
		unit3(a);
		unit3(b);
		cross3(a, b, c);
		double *Oi = O + 9 * i;
		memcpy(Oi, a, sizeof a);
		memcpy(Oi + 3, b, sizeof b);
		memcpy(Oi + 6, c, sizeof c);
		orc_gram_schmidt(Oi);
		memcpy(r + 3 * i, p, sizeof p);
		i++;
	}
	fclose(f);
	return i == n ? n : -1;
}

static int rrr_write(const char *path, long long step, int n, const double box[3], const double *r,
                     const double *O, int generator_header) {
	FILE *f = fopen(path, "w");
	if(!f) return -1;
	if(generator_header) fprintf(f, "0 %d %lf %lf %lf\n", n, box[0], box[1], box[2]);
	else fprintf(f, "%lld %d %.12lf %.12lf %.12lf\n", step, n, box[0], box[1], box[2]);
	for(int i = 0; i < n; i++) {
		const double *Oi = O + 9 * i, *ri = r + 3 * i;
		fprintf(f, "%lf %lf %lf\n", Oi[0], Oi[1], Oi[2]);
		fprintf(f, "%lf %lf %lf\n", Oi[3], Oi[4], Oi[5]);
		fprintf(f, "%.12lf %.12lf %.12lf\n", ri[0], ri[1], ri[2]);
	}
	fclose(f);
	return 0;
}

int orc_rrr_save(const char *path, long long step, int n, const double box[3], const double *r, const double *O) {
	return rrr_write(path, step, n, box, r, O, 0);
}

/* generated.rrr header (generator.c:73): box printed with %lf */
int orc_rrr_save_generated(const char *path, int n, const double box[3], const double *r, const double *O) {
	return rrr_write(path, 0, n, box, r, O, 1);
}

/* ---------------------------------------------------------------------- */
/* Philox4x32-10                                                           */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	uint32_t k0 = key[0], k1 = key[1];
	for(int round = 0; round < 10; round++) {
		uint64_t p0 = (uint64_t) 0xD2511F53u * c0;
		uint64_t p1 = (uint64_t) 0xCD9E8D57u * c2;
		uint32_t n0 = (uint32_t) (p1 >> 32) ^ c1 ^ k0;
		uint32_t n1 = (uint32_t) p1;
		uint32_t n2 = (uint32_t) (p0 >> 32) ^ c3 ^ k1;
		uint32_t n3 = (uint32_t) p0;
		c0 = n0;
		c1 = n1;
		c2 = n2;
		c3 = n3;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

void orc_philox4x32_10_n(int n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
	for(int i = 0; i < n; i++) orc_philox4x32_10(ctr + 4 * (size_t) i, key + 2 * (size_t) i, out + 4 * (size_t) i);
}
