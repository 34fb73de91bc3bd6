/*
 * pp_rrr.c -- .rrr reading and writing with the reference's arithmetic and
 * number formats.  Compiled without contraction (-ffp-contract=off) so each
 * product and sum rounds where it is written.
 */
#include "pp_rrr.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* v / |v| component by component (utils.c:31-36) */
static void unit3(double *v) {
	const double len = sqrt(dot3(v, v));
	v[0] /= len;
	v[1] /= len;
	v[2] /= len;
}

void pprrr_gram_schmidt(double O[9]) {
	double *a = O, *b = O + 3, *c = O + 6;
	const double aa = dot3(a, a);
	const double f_ba = dot3(b, a) / aa;
	for(int k = 0; k < 3; k++) b[k] -= f_ba * a[k];
	const double ca = dot3(c, a), cb = dot3(c, b), bb = dot3(b, b);
	const double f_ca = ca / aa, f_cb = cb / bb;
	for(int k = 0; k < 3; k++) c[k] -= f_ca * a[k] + f_cb * b[k];
	unit3(a);
	unit3(b);
	unit3(c);
}

int pprrr_read_step(const char *path, long long *step) {
	FILE *in = fopen(path, "r");
	if(!in) return -1;
	const int ok = fscanf(in, "%lld %*d %*f %*f %*f\n", step) == 1;
	fclose(in);
	return ok ? 0 : -1;
}

int pprrr_read(const char *path, pprrr *out, char *err, size_t errcap) {
	memset(out, 0, sizeof(*out));
	FILE *in = fopen(path, "r");
	if(!in) {
		snprintf(err, errcap, "Initial_conditions_file '%s' is not readable\n", path);
		return -1;
	}
	if(fscanf(in, "%lld %d %lf %lf %lf\n", &out->step, &out->n, out->box, out->box + 1, out->box + 2) != 5 ||
	   out->n < 0) {
		fclose(in);
		snprintf(err, errcap, "The initial configuration file '%s' is empty or its headers are malformed\n", path);
		return -1;
	}
	const int n = out->n;
	out->r = malloc(sizeof(double) * 3 * (size_t) (n > 0 ? n : 1));
	out->O = malloc(sizeof(double) * 9 * (size_t) (n > 0 ? n : 1));
	if(!out->r || !out->O) {
		fclose(in);
		pprrr_free(out);
		snprintf(err, errcap, "out of memory reading '%s'\n", path);
		return -1;
	}
	/* one particle per group of three rows, as long as a further line exists
	 * (system.c:213-245): row 0 by line, row 1 and the position by fscanf */
	int i = 0;
	char line[512];
	while(fgets(line, sizeof line, in) != NULL) {
		double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, p[3] = {0, 0, 0};
		sscanf(line, "%lf %lf %lf\n", a, a + 1, a + 2);
		if(fscanf(in, "%lf %lf %lf\n", b, b + 1, b + 2) < 0) b[0] = b[1] = b[2] = 0.;
		if(fscanf(in, "%lf %lf %lf\n", p, p + 1, p + 2) < 0) p[0] = p[1] = p[2] = 0.;
		if(i < n) {
			unit3(a);
			unit3(b);
			double *O = out->O + 9 * (size_t) i;
			memcpy(O, a, sizeof a);
			memcpy(O + 3, b, sizeof b);
			O[6] = a[1] * b[2] - a[2] * b[1];
			O[7] = a[2] * b[0] - a[0] * b[2];
			O[8] = a[0] * b[1] - a[1] * b[0];
			pprrr_gram_schmidt(O);
			memcpy(out->r + 3 * (size_t) i, p, sizeof p);
		}
		i++;
	}
	fclose(in);
	if(i != n) {
		snprintf(err, errcap,
		         "Number of particles found in configuration (%d) is different from the value found in the header (%d)\n",
		         i, n);
		pprrr_free(out);
		return -1;
	}
	return 0;
}

void pprrr_free(pprrr *c) {
	free(c->r);
	free(c->O);
	c->r = c->O = NULL;
}

int pprrr_write(const char *path, long long step, int n, const double box[3], const double *r, const double *O) {
	FILE *f = fopen(path, "w");
	if(!f) return -1;
	fprintf(f, "%lld %d %.12lf %.12lf %.12lf\n", step, n, box[0], box[1], box[2]);
	for(int i = 0; i < n; i++) {
		const double *o = O + 9 * (size_t) i, *p = r + 3 * (size_t) i;
		fprintf(f, "%lf %lf %lf\n", o[0], o[1], o[2]);
		fprintf(f, "%lf %lf %lf\n", o[3], o[4], o[5]);
		fprintf(f, "%.12lf %.12lf %.12lf\n", p[0], p[1], p[2]);
	}
	fclose(f);
	return 0;
}
