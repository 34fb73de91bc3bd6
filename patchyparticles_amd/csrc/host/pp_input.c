/*
 * pp_input.c -- "key = value" input files with the reference's semantics
 * (parse_input.c:3-160): blank and '#' lines skipped, a line without '=' is
 * reported and ignored, the first occurrence of a key with a non-empty value
 * wins.
 */
#define _GNU_SOURCE
#include "pp_input.h"

#include <ctype.h>
#include <stdlib.h>
#include <string.h>

typedef struct ppi_entry {
	char *key;
	char *value;
	int reads;
} ppi_entry;

struct ppi_file {
	ppi_entry *e;
	int n, cap;
};

/* copy of s[0, len) without leading and trailing white space */
static char *trimmed(const char *s, size_t len) {
	while(len > 0 && isspace((unsigned char) *s)) {
		s++;
		len--;
	}
	while(len > 0 && isspace((unsigned char) s[len - 1])) len--;
	char *out = malloc(len + 1);
	if(!out) return NULL;
	memcpy(out, s, len);
	out[len] = '\0';
	return out;
}

static int add_entry(ppi_file *f, char *key, char *value) {
	if(f->n == f->cap) {
		int cap = f->cap ? 2 * f->cap : 16;
		ppi_entry *e = realloc(f->e, (size_t) cap * sizeof(ppi_entry));
		if(!e) return -1;
		f->e = e;
		f->cap = cap;
	}
	f->e[f->n].key = key;
	f->e[f->n].value = value;
	f->e[f->n].reads = 0;
	f->n++;
	return 0;
}

ppi_file *ppi_load(const char *path) {
	FILE *in = fopen(path, "r");
	if(!in) {
		fprintf(stderr, "Input file not found\n");
		return NULL;
	}
	ppi_file *f = calloc(1, sizeof(ppi_file));
	char *line = NULL;
	size_t cap = 0;
	ssize_t len;
	while(f && (len = getline(&line, &cap, in)) != -1) {
		if(len > 0 && line[len - 1] == '\n') line[--len] = '\0';
		char *t = trimmed(line, (size_t) len);
		if(!t) break;
		if(t[0] == '\0' || t[0] == '#') {
			free(t);
			continue;
		}
		char *eq = strchr(t, '=');
		if(!eq) {
			fprintf(stderr, "WARNING: malformed line '%s'\n", line);
			free(t);
			continue;
		}
		char *key = trimmed(t, (size_t) (eq - t));
		char *value = trimmed(eq + 1, strlen(eq + 1));
		free(t);
		if(!key || !value || add_entry(f, key, value) != 0) {
			free(key);
			free(value);
			break;
		}
	}
	free(line);
	fclose(in);
	return f;
}

void ppi_free(ppi_file *f) {
	if(!f) return;
	for(int i = 0; i < f->n; i++) {
		free(f->e[i].key);
		free(f->e[i].value);
	}
	free(f->e);
	free(f);
}

static const char *lookup(ppi_file *f, const char *key, int mandatory) {
	for(int i = 0; i < f->n; i++)
		if(strcmp(f->e[i].key, key) == 0 && f->e[i].value[0] != '\0') {
			f->e[i].reads++;
			return f->e[i].value;
		}
	if(mandatory) {
		fprintf(stderr, "Mandatory key '%s' not found, exiting\n", key);
		exit(-1);
	}
	fprintf(stderr, "INFO: optional key '%s' not found\n", key);
	return NULL;
}

int ppi_string(ppi_file *f, const char *key, char *dst, size_t cap, int mandatory) {
	const char *v = lookup(f, key, mandatory);
	if(!v) return PPI_MISSING;
	if(cap > 0) {
		strncpy(dst, v, cap - 1);
		dst[cap - 1] = '\0';
	}
	return PPI_FOUND;
}

int ppi_int(ppi_file *f, const char *key, int *dst, int mandatory) {
	const char *v = lookup(f, key, mandatory);
	if(!v) return PPI_MISSING;
	*dst = atoi(v);
	return PPI_FOUND;
}

int ppi_llint(ppi_file *f, const char *key, long long *dst, int mandatory) {
	const char *v = lookup(f, key, mandatory);
	if(!v) return PPI_MISSING;
	*dst = atol(v);
	return PPI_FOUND;
}

int ppi_double(ppi_file *f, const char *key, double *dst, int mandatory) {
	const char *v = lookup(f, key, mandatory);
	if(!v) return PPI_MISSING;
	*dst = atof(v);
	return PPI_FOUND;
}

void ppi_report_unread(const ppi_file *f, FILE *out) {
	for(int i = 0; i < f->n; i++)
		if(f->e[i].reads == 0) fprintf(out, "WARNING: key '%s' was not used\n", f->e[i].key);
}
