/*
 * ref_shim.c -- a thin driver compiled TOGETHER WITH the reference's own
 * sources (from /root/reference, by oracle/Makefile, into oracle/_ref/), so
 * that tests can call the reference's real functions on chosen inputs.
 *
 * TEST INFRASTRUCTURE ONLY.  It contains no reference code: it includes the
 * reference headers and calls system_init / MC_init / MC_energy /
 * do_ensemble / MC_check_energy / cells_* exactly as main.c does.
 */
#include "MC.h"
#include "cells.h"
#include "defs.h"
#include "output.h"
#include "parse_input.h"
#include "system.h"
#include "utils.h"

#include <stdlib.h>
#include <string.h>

static System S;
static Output IO;
static input_file INP;
static int loaded = 0;

/* system.c / main.c path: parse the input file, load the configuration it
 * names, seed, build cells and set the dynamics (main.c:48-61).
 * Returns N, or -1 if the initial configuration overlaps. */
int ref_load(const char *input_path) {
	memset(&S, 0, sizeof S);
	memset(&IO, 0, sizeof IO);
	IO.log = stderr;
	loadInputFile(&INP, input_path);
	if(INP.state == ERROR) return -2;
	system_init(&INP, &S, &IO);
	MC_init(&INP, &S, &IO);
	S.energy = 0;
	for(int i = 0; i < S.N; i++) {
		S.energy += MC_energy(&S, S.particles + i);
		if(S.overlap == 1) return -1;
	}
	S.energy *= 0.5;
	loaded = 1;
	return S.N;
}

/* Replace positions and orientation rows in memory, recompute the patches
 * (set_patches, utils.c:20-23) and rebuild the cells (cells.c:17-63). */
int ref_set_state(const double *r, const double *O, const double box[3]) {
	if(!loaded) return -1;
	if(box) memcpy(S.box, box, sizeof(double) * 3);
	for(int i = 0; i < S.N; i++) {
		PatchyParticle *p = S.particles + i;
		memcpy(p->r, r + 3 * i, sizeof(double) * 3);
		for(int k = 0; k < 3; k++) memcpy(p->orientation[k], O + 9 * i + 3 * k, sizeof(double) * 3);
		set_patches(&S, p);
	}
	cells_free(S.cells);
	cells_init(&S, &IO, S.r_cut);
	cells_fill(&S);
	return 0;
}

int ref_get_state(double *r, double *O, double *box) {
	if(!loaded) return -1;
	for(int i = 0; i < S.N; i++) {
		PatchyParticle *p = S.particles + i;
		if(r) memcpy(r + 3 * i, p->r, sizeof(double) * 3);
		if(O)
			for(int k = 0; k < 3; k++) memcpy(O + 9 * i + 3 * k, p->orientation[k], sizeof(double) * 3);
	}
	if(box) memcpy(box, S.box, sizeof(double) * 3);
	return S.N;
}

/* Per-particle MC_energy on the current state: bonds[i] = -E_i, overlap[i]. */
int ref_particle_energies(int *bonds, unsigned char *overlap) {
	if(!loaded) return -1;
	for(int i = 0; i < S.N; i++) {
		double e = MC_energy(&S, S.particles + i);
		bonds[i] = (int) (-e);
		overlap[i] = (unsigned char) S.overlap;
	}
	S.overlap = 0;
	return 0;
}

/* The main loop cadence (main.c:78-120) without file output: every
 * print_every steps (except step 0) MC_check_energy and cells_check; then
 * one do_ensemble per step. */
int ref_run(long long steps, long long print_every) {
	if(!loaded) return -1;
	for(long long st = 0; st < steps; st++) {
		if(print_every > 0 && st % print_every == 0) {
			utils_reset_acceptance_counters(&S);
			if(st != 0) {
				MC_check_energy(&S, &IO);
				cells_check(&S, &IO);
			}
		}
		S.do_ensemble(&S, &IO);
	}
	return 0;
}

double ref_energy(void) { return S.energy; }
double ref_volume(void) { return S.V; }

void ref_counters(long long tries[2], long long accepted[2]) {
	tries[0] = S.tries[ROTO_TRASL];
	tries[1] = S.tries[VOLUME];
	accepted[0] = S.accepted[ROTO_TRASL];
	accepted[1] = S.accepted[VOLUME];
}

/* MC_would_interact between particle 0 (moved to rp/Op) and a virtual
 * particle at rq with orientation Oq; uses particle 0 and 1 as scratch.
 * Needs a loaded system with N >= 2. */
int ref_pair(const double rp[3], const double Op[9], const double rq[3], const double Oq[9], const double box[3]) {
	if(!loaded || S.N < 2) return -9;
	memcpy(S.box, box, sizeof(double) * 3);
	PatchyParticle *p = S.particles, *q = S.particles + 1;
	memcpy(p->r, rp, sizeof(double) * 3);
	memcpy(q->r, rq, sizeof(double) * 3);
	for(int k = 0; k < 3; k++) {
		memcpy(p->orientation[k], Op + 3 * k, sizeof(double) * 3);
		memcpy(q->orientation[k], Oq + 3 * k, sizeof(double) * 3);
	}
	set_patches(&S, p);
	set_patches(&S, q);
	int onp = -1, onq = -1;
	return MC_interact(&S, p, q, &onp, &onq);
}

void ref_base_patches(double out[12]) {
	for(int i = 0; i < 4; i++) memcpy(out + 3 * i, S.base_patches[i], sizeof(double) * 3);
}
