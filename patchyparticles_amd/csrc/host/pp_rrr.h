/*
 * pp_rrr.h -- the reference's .rrr configuration format.
 *
 * Header "step N Lx Ly Lz"; then three lines per particle: orientation row 0,
 * orientation row 1, position (output.c:302-317).  Loading normalises rows 0
 * and 1, sets row 2 = row 0 x row 1 and orthonormalises the three rows by
 * Gram-Schmidt (system.c:213-245, utils.c:25-36, 91-113), with every
 * operation in the reference's order so the doubles are the reference's.
 */
#ifndef PP_RRR_H
#define PP_RRR_H

#include <stddef.h>

typedef struct pprrr {
	long long step;
	int n;
	double box[3];
	double *r; /* 3n, particle-index order */
	double *O; /* 9n, orientation rows */
} pprrr;

/* 0 on success; otherwise -1 and a message in err (the reference's wording). */
int pprrr_read(const char *path, pprrr *out, char *err, size_t errcap);
void pprrr_free(pprrr *c);

/* Only the step of the header (the first value of the first row). */
int pprrr_read_step(const char *path, long long *step);

/* 0 on success, -1 if the file cannot be written. */
int pprrr_write(const char *path, long long step, int n, const double box[3], const double *r, const double *O);

/* Gram-Schmidt on the rows of O, each normalised afterwards (utils.c:91-113). */
void pprrr_gram_schmidt(double O[9]);

#endif
