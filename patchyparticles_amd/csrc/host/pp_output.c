/*
 * pp_output.c -- output files with the reference's names, modes and number
 * formats (output.c:19-201, 355-388).
 */
#include "pp_output.h"

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
	if(o->npt) {
		strcpy(name, "density.dat");
		ppi_string(in, "Density_file", name, sizeof name, 0);
		o->density = open_or_die(o, "Density file", name, mode);
	}
	int bonds = 0;
	if(ppi_int(in, "Print_bonds", &bonds, 0) == PPI_FOUND && bonds)
		ppo_log(o, "WARNING: Print_bonds is not supported by this program\n");
}

void ppo_free(ppo *o) {
	if(o->energy) fclose(o->energy);
	if(o->acc) fclose(o->acc);
	if(o->density) fclose(o->density);
	if(o->log && o->log != stderr) fclose(o->log);
	memset(o, 0, sizeof(*o));
}

void ppo_print(ppo *o, long long step, int n, double energy, double volume, long long acc_roto, long long try_roto,
               long long acc_vol, long long try_vol) {
	const double e = n > 0 ? energy / n : 0.;
	fprintf(o->energy, "%lld %lf\n", step, e);
	fflush(o->energy);
	if(o->npt) {
		fprintf(o->density, "%lld %lf %d\n", step, n / volume, n);
		fflush(o->density);
	}
	fprintf(o->acc, "%lld", step);
	fprintf(o->acc, " %e", acc_roto / (double) try_roto);
	if(o->npt) fprintf(o->acc, " %e", acc_vol / (double) try_vol);
	fprintf(o->acc, "\n");
	fflush(o->acc);
	printf("%lld %lf %lf", step, e, energy);
	if(o->npt) printf(" %lf", n / volume);
	printf("\n");
	fflush(stdout);
}
