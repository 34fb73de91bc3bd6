/*
 * pp_output.h -- the reference's output files and log (output.c:19-201,
 * 355-388): energy.dat, acceptance.dat, density.dat (NPT), the configuration
 * folder with conf_<step>.rrr and the "last" configuration, and the line
 * printed on stdout at every print step.
 */
#ifndef PP_OUTPUT_H
#define PP_OUTPUT_H

#include <stdio.h>

#include "pp_input.h"

typedef struct ppo {
	FILE *log;      /* stderr unless Log_file names a file */
	FILE *energy;   /* Energy_file (energy.dat) */
	FILE *acc;      /* Acceptance_file (acceptance.dat) */
	FILE *density;  /* Density_file (density.dat), NPT only */
	long long start_from;   /* first step: 0, or the configuration's step if Restart_step_counter = 0 */
	long long print_every;  /* Print_every */
	long long save_every;   /* Save_every */
	char folder[512];       /* Configuration_folder */
	char last[512];         /* Configuration_last (last.rrr) */
	int npt;                /* a non-canonical ensemble (NPT or GC): density.dat is written */
	int gc;                 /* grand canonical: insertion and removal acceptances */
	int print_bonds;        /* Print_bonds (NVT only) */
	int lx;                 /* box-shape moves (set by the driver, NVT only) */
	FILE *boxshape;         /* boxshape.dat, opened when Lx_move is set (output.c:101-108) */
	char bonds_folder[512]; /* Bonds_folder */
} ppo;

/* Reads the output keys and opens the files; exits with a message on error. */
void ppo_init(ppo *o, ppi_file *in, int ensemble, const char *initial_conf);
void ppo_free(ppo *o);

/* A message on the log (and on stderr when the log is a file). */
void ppo_log(ppo *o, const char *fmt, ...);
/* ppo_log, then exit(1) (output.c:374-388). */
void ppo_exit(ppo *o, const char *fmt, ...);

/* The print-step lines: energy.dat "step E/N", density.dat "step N/V N"
 * (NPT, GC), boxshape.dat "step Lx Ly Lz" (box-shape moves), acceptance.dat
 * "step acc_roto [acc_add acc_remove | acc_volume] [acc_Lx]", stdout
 * "step E/N E [N/V]" (output.c:119-201).  Acceptances are accepted / tries
 * since the previous print; acc_vol / try_vol count the volume moves (NPT) or
 * the box-shape moves (NVT, GC); gc_acc / gc_try the insertions and the
 * removals (GC). */
void ppo_print(ppo *o, long long step, int n, double energy, double volume, const double box[3], long long acc_roto,
               long long try_roto, long long acc_vol, long long try_vol, const long long gc_acc[2],
               const long long gc_try[2]);

/* <Bonds_folder>/bonds_<step>.dat (output.c:205-268): N, then per particle
 * "index count" and the line of its bonded partners, each followed by a
 * space.  count[i] partners of particle i are at partners[i * max_per + k];
 * they are listed in the reference's order for a freshly filled cell list:
 * the 27 cells around i on the reference grid (side >= 1 + kf_delta, at least
 * 3 per side, cells.c:17-49) in loop order, within a cell by decreasing index
 * (cells_fill pushes to the front, cells.c:51-62). */
void ppo_print_bonds(ppo *o, long long step, int n, const double box[3], const double *r, double kf_delta,
                     const int *count, const int *partners, int max_per);

#endif
