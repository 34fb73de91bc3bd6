/*
 * pp_output.c -- output files with the reference's names, modes and number
 * formats (output.c:19-201, 355-388).
 */
#include "pp_output.h"

#include <float.h>
#include <math.h>

#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "pp_rrr.h"

static void vlog(ppo *o, const char *fmt, va_list ap) {
	va_list cp;
	if(o->log && o->log != stderr) {
		va_copy(cp, ap);
		vfprintf(o->log, fmt, cp);
		va_end(cp);
		fflush(o->log);
	}
	vfprintf(stderr, fmt, ap);
}

void ppo_log(ppo *o, const char *fmt, ...) {
	va_list ap;
	va_start(ap, fmt);
	vlog(o, fmt, ap);
	va_end(ap);
}

void ppo_exit(ppo *o, const char *fmt, ...) {
	va_list ap;
	va_start(ap, fmt);
	vlog(o, fmt, ap);
	va_end(ap);
	exit(1);
}

static FILE *open_or_die(ppo *o, const char *what, const char *name, const char *mode) {
	FILE *f = fopen(name, mode);
	if(!f) ppo_exit(o, "%s '%s' is not writable\n", what, name);
	return f;
}

void ppo_init(ppo *o, ppi_file *in, int ensemble, const char *initial_conf) {
	memset(o, 0, sizeof(*o));
	o->log = stderr;
	int restart = 1;
	ppi_int(in, "Restart_step_counter", &restart, 0);
	/* a fresh step counter truncates the output files, a continued one appends */
	const char *mode = restart ? "w" : "a";

	char name[512];
	if(ppi_string(in, "Log_file", name, sizeof name, 0) == PPI_FOUND && strcmp(name, "none") != 0) {
		o->log = fopen(name, "a");
		if(!o->log) {
			o->log = stderr;
			ppo_exit(o, "Log file '%s' is not writable\n", name);
		}
	}
	if(!restart) {
		if(pprrr_read_step(initial_conf, &o->start_from) != 0)
			ppo_exit(o, "Invalid initial configuration: the first value in the first row should be the time step "
			            "of the configuration");
	}
	ppi_llint(in, "Print_every", &o->print_every, 1);
	ppi_llint(in, "Save_every", &o->save_every, 1);
	if(o->print_every <= 0 || o->save_every <= 0) ppo_exit(o, "Print_every and Save_every must be positive\n");

	int mgl = 0;
	if(ppi_int(in, "Save_also_as_mgl", &mgl, 0) == PPI_FOUND && mgl)
		ppo_log(o, "WARNING: Save_also_as_mgl is not supported by this program; only .rrr files are written\n");

	strcpy(name, "energy.dat");
	ppi_string(in, "Energy_file", name, sizeof name, 0);
	o->energy = open_or_die(o, "Energy file", name, mode);
	strcpy(name, "acceptance.dat");
	ppi_string(in, "Acceptance_file", name, sizeof name, 0);
	o->acc = open_or_die(o, "Acceptance file", name, mode);

	ppi_string(in, "Configuration_folder", o->folder, sizeof o->folder, 1);
	if(access(o->folder, W_OK) != 0)
		ppo_exit(o, "Cannot create files in directory '%s': please make sure that the directory exists and it is "
		            "writable\n",
		         o->folder);
	strcpy(o->last, "last.rrr");
	ppi_string(in, "Configuration_last", o->last, sizeof o->last, 0);

	o->npt = ensemble != 0;
	o->gc = ensemble == 1;
	if(o->npt) {
		strcpy(name, "density.dat");
		ppi_string(in, "Density_file", name, sizeof name, 0);
		o->density = open_or_die(o, "Density file", name, mode);
	}
	o->print_bonds = 0;
	ppi_int(in, "Print_bonds", &o->print_bonds, 0);
	if(o->print_bonds) {
		if(o->npt) ppo_exit(o, "The printing of bonds is unavailable in non-canonical ensembles");
		ppi_string(in, "Bonds_folder", o->bonds_folder, sizeof o->bonds_folder, 1);
		if(access(o->bonds_folder, W_OK) != 0)
			ppo_exit(o, "Cannot create files in directory '%s': please make sure that the directory exists and it is "
			            "writable\n",
			         o->bonds_folder);
	}
	int shape = 0;
	ppi_int(in, "Lx_move", &shape, 0);
	if(shape) {
		o->boxshape = fopen("boxshape.dat", mode);
		if(!o->boxshape) ppo_exit(o, "boxshape.dat is not writable\n");
	}
}

/* the reference's cell of a position (cells_fill_and_get_idx_from_vector) */
static void ref_cell(const double r[3], const double box[3], const int ns[3], int idx[3]) {
	for(int k = 0; k < 3; k++) idx[k] = (int) ((r[k] / box[k] - floor(r[k] / box[k])) * (1. - DBL_EPSILON) * ns[k]);
}

/* loop offset (-1, 0, 1) of cell b from cell a along an axis of ns cells */
static int cell_offset(int a, int b, int ns) {
	for(int j = -1; j < 2; j++)
		if((a + j + ns) % ns == b) return j;
	return 2; /* not a neighbour cell: listed last */
}

void ppo_print_bonds(ppo *o, long long step, int n, const double box[3], const double *r, double kf_delta,
                     const int *count, const int *partners, int max_per) {
	char name[1100];
	snprintf(name, sizeof name, "%s/bonds_%lld.dat", o->bonds_folder, step);
	FILE *out = fopen(name, "w");
	if(!out) ppo_exit(o, "File '%s' is not writable\n", name);
	const double rcut = 1. + kf_delta;
	int ns[3];
	for(int k = 0; k < 3; k++) {
		ns[k] = (int) floor(box[k] / rcut);
		if(ns[k] < 3) ns[k] = 3;
	}
	fprintf(out, "%d\n", n);
	long long *key = malloc(sizeof(long long) * (size_t) (max_per > 0 ? max_per : 1));
	int *lst = malloc(sizeof(int) * (size_t) (max_per > 0 ? max_per : 1));
	if(!key || !lst) ppo_exit(o, "out of memory\n");
	for(int i = 0; i < n; i++) {
		const int m = count[i] < max_per ? count[i] : max_per;
		int ci[3];
		ref_cell(r + 3 * i, box, ns, ci);
		for(int k = 0; k < m; k++) {
			const int q = partners[(size_t) i * max_per + k];
			int cq[3];
			ref_cell(r + 3 * q, box, ns, cq);
			const int o0 = cell_offset(ci[0], cq[0], ns[0]), o1 = cell_offset(ci[1], cq[1], ns[1]),
			          o2 = cell_offset(ci[2], cq[2], ns[2]);
			/* loop order over the 27 cells, then decreasing index */
			key[k] = ((long long) ((o0 + 1) * 16 + (o1 + 1) * 4 + (o2 + 1)) << 32) | (unsigned) (0x7fffffff - q);
			lst[k] = q;
		}
		for(int a = 1; a < m; a++) /* insertion sort: a handful of partners */
			for(int b = a; b > 0 && key[b - 1] > key[b]; b--) {
				const long long tk = key[b];
				key[b] = key[b - 1];
				key[b - 1] = tk;
				const int tq = lst[b];
				lst[b] = lst[b - 1];
				lst[b - 1] = tq;
			}
		fprintf(out, "%d %d\n", i, count[i]);
		for(int k = 0; k < m; k++) fprintf(out, "%d ", lst[k]);
		fprintf(out, "\n");
	}
	free(key);
	free(lst);
	fclose(out);
}

void ppo_free(ppo *o) {
	if(o->energy) fclose(o->energy);
	if(o->acc) fclose(o->acc);
	if(o->density) fclose(o->density);
	if(o->boxshape) fclose(o->boxshape);
	if(o->log && o->log != stderr) fclose(o->log);
	memset(o, 0, sizeof(*o));
}

void ppo_print(ppo *o, long long step, int n, double energy, double volume, const double box[3], long long acc_roto,
               long long try_roto, long long acc_vol, long long try_vol, const long long gc_acc[2],
               const long long gc_try[2]) {
	const double e = n > 0 ? energy / n : 0.;
	fprintf(o->energy, "%lld %lf\n", step, e);
	fflush(o->energy);
	if(o->npt) {
		fprintf(o->density, "%lld %lf %d\n", step, n / volume, n);
		fflush(o->density);
	}
	if(o->lx && o->boxshape) {
		fprintf(o->boxshape, "%lld %lf %lf %lf\n", step, box[0], box[1], box[2]);
		fflush(o->boxshape);
	}
	fprintf(o->acc, "%lld", step);
	fprintf(o->acc, " %e", acc_roto / (double) try_roto);
	if(o->gc) {
		/* ADD, REMOVE (output.c:166-169), then the box-shape moves */
		fprintf(o->acc, " %e", gc_acc[0] / (double) gc_try[0]);
		fprintf(o->acc, " %e", gc_acc[1] / (double) gc_try[1]);
		if(o->lx) fprintf(o->acc, " %e", acc_vol / (double) try_vol);
	}
	else if(o->npt || o->lx) fprintf(o->acc, " %e", acc_vol / (double) try_vol);
	fprintf(o->acc, "\n");
	fflush(o->acc);
	printf("%lld %lf %lf", step, e, energy);
	if(o->npt) printf(" %lf", n / volume);
	printf("\n");
	fflush(stdout);
}
