/*
 * pp_input.h -- reader of the reference's input files: one "key = value" per
 * line, '#' starts a comment line, keys and values trimmed of white space
 * (parse_input.c:3-160).  A key counts as present only with a non-empty
 * value; integers convert as atoi/atol, reals as atof, as the reference does.
 */
#ifndef PP_INPUT_H
#define PP_INPUT_H

#include <stddef.h>
#include <stdio.h>

#define PPI_FOUND 0
#define PPI_MISSING (-1)

typedef struct ppi_file ppi_file;

/* NULL (after a message on stderr) if the file cannot be read. */
ppi_file *ppi_load(const char *path);
void ppi_free(ppi_file *f);

/* Each getter returns PPI_FOUND or PPI_MISSING.  A missing mandatory key
 * prints "Mandatory key '<key>' not found, exiting" and exits with status -1;
 * a missing optional key is reported on stderr and leaves *dst unchanged
 * (parse_input.c:100-115). */
int ppi_string(ppi_file *f, const char *key, char *dst, size_t cap, int mandatory);
int ppi_int(ppi_file *f, const char *key, int *dst, int mandatory);
int ppi_llint(ppi_file *f, const char *key, long long *dst, int mandatory);
int ppi_double(ppi_file *f, const char *key, double *dst, int mandatory);

/* Keys of the file that no getter asked for, one line each on out. */
void ppi_report_unread(const ppi_file *f, FILE *out);

#endif
