/*
 * pp_generator.c -- generator_hip: the reference's generator program
 * (generator.c:86-124).  "generator_hip N density [seed]" writes
 * generated.rrr: N non-overlapping unit spheres inserted at random in a
 * cubic box of side (N/density)^(1/3), uniformly random orientations.
 * Without a seed the insertion stream is the reference generator's (the
 * 48-bit LCG in its unseeded state), so the file matches its output byte for
 * byte; a seed starts the stream as srand48(seed).  Header and rows use the
 * generator's own number formats (generator.c:70-81).
 */
#include <stdio.h>
#include <stdlib.h>

#include "pp.h"

int main(int argc, char *argv[]) {
	if(argc < 3) {
		fprintf(stderr, "Usage is %s N density\n", argv[0]);
		return 1;
	}
	const int n = atoi(argv[1]);
	const double density = atof(argv[2]);
	const long seed = argc > 3 ? atol(argv[3]) : -1;
	if(density > 0.7) {
		fprintf(stderr, "It is very hard to generate very dense configurations by just randomly inserting particles. "
		                "This simple generator cannot produce configurations with density higher than 0.7\n");
		return 1;
	}
	if(n < 0 || !(density > 0.)) {
		fprintf(stderr, "N must be >= 0 and the density > 0\n");
		return 1;
	}
	double *r = malloc(sizeof(double) * 3 * (size_t) (n > 0 ? n : 1));
	double *O = malloc(sizeof(double) * 9 * (size_t) (n > 0 ? n : 1));
	double box[3];
	if(!r || !O) {
		fprintf(stderr, "out of memory\n");
		return 1;
	}
	const pp_status st = pp_generate(n, density, seed, r, O, box);
	if(st != PP_OK) {
		fprintf(stderr, "pp_generate failed: %s\n", pp_strerror(st));
		return 1;
	}
	const char *name = "generated.rrr";
	FILE *out = fopen(name, "w");
	if(!out) {
		fprintf(stderr, "File '%s' is not writable\n", name);
		return 1;
	}
	fprintf(out, "0 %d %lf %lf %lf\n", n, box[0], box[1], box[2]);
	for(int i = 0; i < n; i++) {
		const double *o = O + 9 * (size_t) i, *p = r + 3 * (size_t) i;
		fprintf(out, "%lf %lf %lf\n", o[0], o[1], o[2]);
		fprintf(out, "%lf %lf %lf\n", o[3], o[4], o[5]);
		fprintf(out, "%.12lf %.12lf %.12lf\n", p[0], p[1], p[2]);
	}
	fclose(out);
	fprintf(stderr, "Generation done. The new configuration has been written to the file '%s'\n", name);
	free(r);
	free(O);
	return 0;
}
