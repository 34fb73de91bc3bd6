/*
 * cpu_bench.c -- times the oracle's serial reference-algorithm sampler
 * (kf_oracle.c) on one core: the "port" CPU baseline of bench.py.
 *
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY.
 *
 * usage: cpu_bench N rho T seed max_seconds [min_sweeps]
 * Start: random sequential insertion (orc_generate) with the given seed,
 * uniform random orientations.  NVT, Disp_max 0.05, Theta_max 0.1,
 * KF_delta 0.12, KF_cosmax 0.92.  One warm-up sweep, then whole sweeps until
 * max_seconds have passed.  Prints one JSON line.
 */
#include "kf_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <time.h>

static double now(void) {
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv) {
	if(argc < 6) {
		fprintf(stderr, "usage: %s N rho T seed max_seconds [min_sweeps]\n", argv[0]);
		return 1;
	}
	int n = atoi(argv[1]);
	double rho = atof(argv[2]);
	double T = atof(argv[3]);
	long seed = atol(argv[4]);
	double tmax = atof(argv[5]);
	long long min_sweeps = argc > 6 ? atoll(argv[6]) : 1;
	double *r = malloc(sizeof(double) * 3 * n), *O = malloc(sizeof(double) * 9 * n), box[3];
	if(orc_generate(n, rho, seed, r, O, box) != 0) return 2;
	orc_params p = {0};
	p.ensemble = 0;
	p.temperature = T;
	p.disp_max = 0.05;
	p.theta_max = 0.1;
	p.kf_delta = 0.12;
	p.kf_cosmax = 0.92;
	p.seed = seed;
	orc_sim *s = orc_sim_create(&p, n, r, O, box);
	if(orc_sim_init_energy(s)) return 3;
	orc_sim_sweeps(s, 1);
	double t0 = now(), t1 = t0;
	long long done = 0;
	while(done < min_sweeps || t1 - t0 < tmax) {
		orc_sim_sweeps(s, 1);
		done++;
		t1 = now();
	}
	long long tr[2], ac[2];
	orc_sim_counters(s, tr, ac);
	printf("{\"sweeps\": %lld, \"seconds\": %.6f, \"sweeps_per_s\": %.6f, \"E_per_N\": %.6f, \"acceptance\": %.6f}\n",
	       done, t1 - t0, done / (t1 - t0), orc_sim_energy(s) / n, (double) ac[0] / (double) tr[0]);
	orc_sim_destroy(s);
	free(r);
	free(O);
	return 0;
}
