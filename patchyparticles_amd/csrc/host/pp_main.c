/*
 * pp_main.c -- PatchyParticles_hip: the reference's driver program (main.c)
 * for Dynamics = 0 (rototranslation MC) in Ensemble 0 (NVT), 1 (GC, keys
 * Activity and GC_N_max, system.c:90-95) or 5 (NPT), with
 * every sweep run on the GPU through libpp_hip.so.
 *
 * Same input keys (system.c:58-89, output.c:19-110), same .rrr files and the
 * same energy / density / acceptance outputs and stdout lines at the same
 * cadence (main.c:78-125): at every Print_every step the outputs are written,
 * last.rrr is refreshed and, after step 0, the energy is recomputed from
 * scratch and checked (MC_check_energy, cells_check); at every Save_every
 * step (after 0) conf_<step>.rrr and last.rrr are written, from a snapshot
 * copied to the host while the next block of sweeps runs.  Between two such
 * steps the sweeps run on the GPU in one call.  SIGTERM, SIGABRT, SIGINT and
 * SIGUSR2 end the run after the current block of sweeps, with the last
 * configuration saved (main.c:16-24).
 *
 * Keys of this program only (optional): GPU_device (HIP ordinal, 0),
 * Threads_per_cell (16, 32 or 64), Trials_per_cell (0: a sweep is N attempts;
 * s > 0: s trials per non-empty sweep cell and colour pass), Checkpoint_file
 * (written with every last.rrr: the exact state and the random-number
 * counter), Resume (1: continue bit for bit from Checkpoint_file; the
 * Initial_conditions_file then only gives N, normally the last.rrr written
 * with the checkpoint).  Reference keys beyond the ones above: Print_bonds /
 * Bonds_folder (bond lists) and Lx_move / Lx_change_max / Lyz_min / Lyz_max
 * (box-shape moves, boxshape.dat; accepted by Metropolis, with a warning at
 * start-up, or with the reference's own test when this program's key
 * Lx_reference_acceptance = 1).  The learned proposal: Learned_fraction,
 * Learned_weights, Learned_train_steps, Learned_train_batch,
 * Learned_pos_noise (see setup_learned).
 */
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pp.h"
#include "pp_input.h"
#include "pp_output.h"
#include "pp_rrr.h"

static volatile sig_atomic_t g_stop = 0;

static void on_signal(int sig) {
	signal(sig, SIG_DFL); /* a second signal ends the program at once */
	g_stop = 1;
}

/* sweeps per library call between two output steps: keeps a signal answered
 * within a fraction of a second at any size */
#define MAX_BLOCK 2000

typedef struct run {
	ppo out;
	pp_ctx *ctx;
	pp_net *net; /* learned proposal (NULL: plain moves only) */
	int n;
	int npt;
	double *r, *O;
	char chk[1024];  /* checkpoint path ("" : none) */
	double kf_delta;
	int *bcount, *bpart; /* bond list of Print_bonds */
} run;

#define MAX_BONDS 16 /* bonded partners listed per particle */

static void check(run *u, pp_status st, const char *what) {
	if(st == PP_OK) return;
	ppo_exit(&u->out, "%s failed: %s (%s)\n", what, pp_strerror(st), pp_last_error());
}

/* particles in the system now (grand canonical: changes with the run) */
static int current_n(run *u) {
	pp_counters c;
	check(u, pp_get_counters(u->ctx, 0, &c), "pp_get_counters");
	return c.n;
}

static void save_conf(run *u, long long step, const char *path) {
	double box[3];
	const int n = current_n(u);
	check(u, pp_download(u->ctx, 0, u->r, u->O, box), "pp_download");
	if(pprrr_write(path, step, n, box, u->r, u->O) != 0) ppo_exit(&u->out, "File '%s' is not writable\n", path);
}

/* A save step's configuration in flight: pp_snapshot_start copies the state
 * on the device and the next block of sweeps runs while it travels to the
 * host; conf_<step>.rrr and last.rrr are written from it before anything else
 * happens at the next output step (or at the end of the run). */
typedef struct pending_save {
	int on;
	long long step;
	int n;
	char name[1100];
} pending_save;

static void save_start(run *u, pending_save *p, long long step, const char *name) {
	p->n = current_n(u);
	check(u, pp_snapshot_start(u->ctx, 0), "pp_snapshot_start");
	p->on = 1;
	p->step = step;
	snprintf(p->name, sizeof p->name, "%s", name);
}

static void save_finish(run *u, pending_save *p) {
	if(!p->on) return;
	double box[3];
	p->on = 0;
	check(u, pp_snapshot_finish(u->ctx, u->r, u->O, box), "pp_snapshot_finish");
	if(pprrr_write(p->name, p->step, p->n, box, u->r, u->O) != 0) ppo_exit(&u->out, "File '%s' is not writable\n", p->name);
	if(pprrr_write(u->out.last, p->step, p->n, box, u->r, u->O) != 0)
		ppo_exit(&u->out, "File '%s' is not writable\n", u->out.last);
}

/* the checkpoint: the state at the top of an output step, before its energy
 * check (whose Gram-Schmidt changes the orientations); a resumed run redoes
 * that step's check, as the run that wrote it did */
static void save_checkpoint(run *u) {
	if(u->chk[0]) check(u, pp_checkpoint_save(u->ctx, u->chk), "pp_checkpoint_save");
}

static void print_step(run *u, long long step) {
	pp_counters c;
	check(u, pp_get_counters(u->ctx, 0, &c), "pp_get_counters");
	ppo_print(&u->out, step, c.n, -(double) c.bonds, c.volume, c.box, c.accepted[0], c.tries[0], c.accepted[1],
	          c.tries[1], c.gc_accepted, c.gc_tries);
	check(u, pp_reset_counters(u->ctx), "pp_reset_counters");
	save_conf(u, step, u->out.last);
	if(u->out.print_bonds) {
		double box[3];
		check(u, pp_download(u->ctx, 0, u->r, NULL, box), "pp_download");
		check(u, pp_bond_list(u->ctx, 0, MAX_BONDS, u->bcount, u->bpart), "pp_bond_list");
		for(int i = 0; i < u->n; i++)
			if(u->bcount[i] > MAX_BONDS) ppo_exit(&u->out, "particle %d has more than %d bonded partners\n", i, MAX_BONDS);
		ppo_print_bonds(&u->out, step, u->n, box, u->r, u->kf_delta, u->bcount, u->bpart, MAX_BONDS);
	}
}

/* MC_check_energy + cells_check (MC.c:766-784, cells.c:65-76) */
static void check_energy(run *u) {
	const pp_status st = pp_check_energy(u->ctx, NULL);
	if(st == PP_ERR_OVERLAP) ppo_exit(&u->out, "\nComputing energy from scratch resulted in an overlap, aborting\n");
	if(st == PP_ERR_ENERGY || st == PP_ERR_CELLS) ppo_exit(&u->out, "\n%s\n", pp_last_error());
	check(u, st, "pp_check_energy");
}

static long long next_multiple(long long x, long long m) { return (x / m + 1) * m; }

/* The learned proposal (keys of this program only): Learned_fraction > 0
 * turns it on for that share of the colour passes.  The network's weights
 * come from Learned_weights if that file exists; otherwise a network is
 * trained on a separate copy of the start (Learned_train_steps Adam steps of
 * Learned_train_batch moves, one plain sweep between steps, displacement
 * noise Learned_pos_noise), so the run's own trajectory is not advanced, and
 * saved to Learned_weights when the key is given. */
static void setup_learned(run *u, ppi_file *in, const pp_params *prm, const pp_options *opt, const pprrr *c) {
	double frac = 0.;
	ppi_double(in, "Learned_fraction", &frac, 0);
	if(!(frac > 0.)) return;
	int steps = 10000, batch = 4096;
	double noise = 0.03;
	char wfile[512] = "";
	ppi_int(in, "Learned_train_steps", &steps, 0);
	ppi_int(in, "Learned_train_batch", &batch, 0);
	ppi_double(in, "Learned_pos_noise", &noise, 0);
	ppi_string(in, "Learned_weights", wfile, sizeof wfile, 0);
	check(u, pp_net_create(opt->device, opt->seed ^ 0x5EEDull, &u->net), "pp_net_create");
	float *w = malloc(sizeof(float) * (PP_NET_WEIGHTS + PP_NET_BIASES));
	if(!w) ppo_exit(&u->out, "out of memory\n");
	FILE *f = wfile[0] ? fopen(wfile, "rb") : NULL;
	if(f) {
		const size_t got = fread(w, sizeof(float), PP_NET_WEIGHTS + PP_NET_BIASES, f);
		fclose(f);
		if(got != PP_NET_WEIGHTS + PP_NET_BIASES) ppo_exit(&u->out, "Learned_weights '%s' is not a network file\n", wfile);
		check(u, pp_net_set(u->net, w, w + PP_NET_WEIGHTS), "pp_net_set");
		ppo_log(&u->out, "Learned proposal: weights read from '%s'\n", wfile);
	}
	else {
		pp_options o2 = *opt;
		o2.seed = opt->seed + 1;
		pp_ctx *tc = NULL;
		check(u, pp_create(prm, &o2, c->n, c->r, c->O, c->box, &tc), "pp_create (training copy)");
		double l0 = 0., l1 = 0.;
		check(u, pp_net_train(tc, u->net, steps, batch, 1e-3, noise, prm->theta_max, 1, &l0, &l1), "pp_net_train");
		pp_destroy(tc);
		ppo_log(&u->out, "Learned proposal: %d training steps, loss %lf -> %lf\n", steps, l0, l1);
		if(wfile[0]) {
			check(u, pp_net_get(u->net, w, w + PP_NET_WEIGHTS), "pp_net_get");
			f = fopen(wfile, "wb");
			if(!f || fwrite(w, sizeof(float), PP_NET_WEIGHTS + PP_NET_BIASES, f) != PP_NET_WEIGHTS + PP_NET_BIASES)
				ppo_exit(&u->out, "Learned_weights '%s' is not writable\n", wfile);
			fclose(f);
		}
	}
	free(w);
	check(u, pp_net_attach(u->ctx, u->net, frac, prm->theta_max), "pp_net_attach");
}

int main(int argc, char *argv[]) {
	if(argc == 1) {
		fprintf(stderr, "Usage is %s input\n", argv[0]);
		return 1;
	}
	signal(SIGTERM, on_signal);
	signal(SIGABRT, on_signal);
	signal(SIGINT, on_signal);
	signal(SIGUSR2, on_signal);

	ppi_file *in = ppi_load(argv[1]);
	if(!in) return 1;
	run u;
	memset(&u, 0, sizeof u);

	pp_params prm;
	memset(&prm, 0, sizeof prm);
	int dynamics = 0;
	char conf[512];
	ppi_int(in, "Ensemble", &prm.ensemble, 1);
	ppi_string(in, "Initial_conditions_file", conf, sizeof conf, 1);
	ppo_init(&u.out, in, prm.ensemble, conf);

	ppi_int(in, "Dynamics", &dynamics, 1);
	if(dynamics != 0) ppo_exit(&u.out, "Dynamics %d not supported\n", dynamics);
	if(prm.ensemble != PP_NVT && prm.ensemble != PP_NPT && prm.ensemble != PP_GC)
		ppo_exit(&u.out, "Ensemble %d not supported\n", prm.ensemble);
	ppi_double(in, "Disp_max", &prm.disp_max, 1);
	ppi_double(in, "Theta_max", &prm.theta_max, 1);
	ppi_double(in, "Temperature", &prm.temperature, 1);

	int seed = 0;
	if(ppi_int(in, "Seed", &seed, 0) == PPI_MISSING) {
		seed = (int) time(NULL);
		ppo_log(&u.out, "Using seed %d\n", seed);
	}

	pprrr c;
	char err[1024];
	if(pprrr_read(conf, &c, err, sizeof err) != 0) ppo_exit(&u.out, "%s", err);
	u.n = c.n;
	u.npt = prm.ensemble != PP_NVT;

	ppi_double(in, "KF_delta", &prm.kf_delta, 1);
	ppi_double(in, "KF_cosmax", &prm.kf_cosmax, 1);
	ppo_log(&u.out, "Patch parameters: cosmax = %lf, delta = %lf\n", prm.kf_cosmax, prm.kf_delta);
	if(prm.ensemble == PP_NPT) {
		ppi_double(in, "rescale_factor_max", &prm.rescale_factor_max, 1);
		ppi_double(in, "P", &prm.pressure, 1);
	}
	if(prm.ensemble == PP_GC) {
		/* system.c:90-95 */
		ppi_double(in, "Activity", &prm.activity, 1);
		ppi_int(in, "GC_N_max", &prm.gc_n_max, 1);
		if(prm.gc_n_max < c.n)
			ppo_exit(&u.out, "GC_N_max (%d) is smaller than the number of particles (%d)\n", prm.gc_n_max, c.n);
		u.n = prm.gc_n_max; /* the arrays hold the most particles the run may have */
	}
	/* box-shape moves (system.c:170-196): NVT only */
	int lx = 0;
	double lx_change_max = 0., lyz_min = 0., lyz_max = 0.;
	if(prm.ensemble != PP_NPT) {
		ppi_int(in, "Lx_move", &lx, 0);
		if(lx) {
			if(c.box[1] != c.box[2]) ppo_exit(&u.out, "Lx_move = 1 requires Ly = Lz\n");
			ppi_double(in, "Lx_change_max", &lx_change_max, 1);
			ppi_double(in, "Lyz_min", &lyz_min, 1);
			ppi_double(in, "Lyz_max", &lyz_max, 1);
			if(lyz_min >= c.box[1]) ppo_exit(&u.out, "Lyz_min should be smaller than box size\n");
			if(lyz_max <= c.box[1]) ppo_exit(&u.out, "Lyz_max should be larger than box size\n");
			ppo_log(&u.out, "Ly and Lz are allowed to vary between %lf and %lf\n", lyz_min, lyz_max);
		}
	}
	int lx_ref = 0;
	if(lx) {
		ppi_int(in, "Lx_reference_acceptance", &lx_ref, 0);
		if(lx_ref)
			ppo_log(&u.out, "Lx moves: accepted with exp(+dE/T) as the reference program does (MC.c:562); this rule "
			                "does not sample the Boltzmann distribution\n");
		else
			ppo_log(&u.out, "WARNING: Lx moves are accepted with min(1, exp(-dE/T)) (Metropolis); the reference program "
			                "accepts them with exp(+dE/T) (MC.c:562). Set Lx_reference_acceptance = 1 for the "
			                "reference's rule\n");
	}
	u.out.lx = lx != 0;

	pp_options opt;
	memset(&opt, 0, sizeof opt);
	opt.seed = (uint64_t) (uint32_t) seed;
	ppi_int(in, "GPU_device", &opt.device, 0);
	ppi_int(in, "Threads_per_cell", &opt.threads_per_cell, 0);
	ppi_int(in, "Trials_per_cell", &opt.trials_per_cell, 0);

	const pp_status st = pp_create(&prm, &opt, c.n, c.r, c.O, c.box, &u.ctx);
	if(st == PP_ERR_OVERLAP) ppo_exit(&u.out, "Initial configuration contains an overlap, aborting\n");
	if(st != PP_OK) ppo_exit(&u.out, "pp_create failed: %s (%s)\n", pp_strerror(st), pp_last_error());
	if(lx) {
		check(&u, pp_set_lx_move(u.ctx, 0, lx_change_max, lyz_min, lyz_max), "pp_set_lx_move");
		check(&u, pp_set_lx_acceptance(u.ctx, 0, lx_ref ? PP_LX_REFERENCE : PP_LX_METROPOLIS), "pp_set_lx_acceptance");
	}
	setup_learned(&u, in, &prm, &opt, &c);
	pprrr_free(&c);
	u.r = malloc(sizeof(double) * 3 * (size_t) (u.n > 0 ? u.n : 1));
	u.O = malloc(sizeof(double) * 9 * (size_t) (u.n > 0 ? u.n : 1));
	u.kf_delta = prm.kf_delta;
	u.bcount = malloc(sizeof(int) * (size_t) (u.n > 0 ? u.n : 1));
	u.bpart = malloc(sizeof(int) * (size_t) MAX_BONDS * (size_t) (u.n > 0 ? u.n : 1));
	if(!u.r || !u.O || !u.bcount || !u.bpart) ppo_exit(&u.out, "out of memory\n");

	int resume = 0;
	u.chk[0] = '\0';
	ppi_string(in, "Checkpoint_file", u.chk, sizeof u.chk, 0);
	ppi_int(in, "Resume", &resume, 0);
	if(resume) {
		if(!u.chk[0]) ppo_exit(&u.out, "Resume = 1 needs Checkpoint_file\n");
		check(&u, pp_checkpoint_load(u.ctx, u.chk), "pp_checkpoint_load");
	}

	pp_counters cnt;
	check(&u, pp_get_counters(u.ctx, 0, &cnt), "pp_get_counters");
	ppo_log(&u.out, "Initial energy: %lf\n", cnt.n > 0 ? -(double) cnt.bonds / cnt.n : 0.);

	long long steps = 0;
	ppi_llint(in, "Steps", &steps, 1);
	steps += u.out.start_from;

	long long step = u.out.start_from;
	pending_save pend = {0, 0, 0, {0}};
	while(step < steps && !g_stop) {
		const int print = step % u.out.print_every == 0, save = step > 0 && step % u.out.save_every == 0;
		save_finish(&u, &pend);
		if(print) {
			print_step(&u, step);
			save_checkpoint(&u);
			if(step != 0) check_energy(&u);
		}
		if(save) {
			char name[1100];
			snprintf(name, sizeof name, "%s/conf_%lld.rrr", u.out.folder, step);
			save_start(&u, &pend, step, name);
			if(!print) save_checkpoint(&u);
		}
		long long until = next_multiple(step, u.out.print_every);
		const long long s2 = next_multiple(step, u.out.save_every);
		if(s2 < until) until = s2;
		if(steps < until) until = steps;
		if(until - step > MAX_BLOCK) until = step + MAX_BLOCK;
		check(&u, pp_sweeps(u.ctx, until - step), "pp_sweeps");
		step = until;
	}
	save_finish(&u, &pend);

	save_conf(&u, step, u.out.last);
	save_checkpoint(&u);
	check(&u, pp_get_counters(u.ctx, 0, &cnt), "pp_get_counters");
	const double E = -(double) cnt.bonds;
	printf("%lld %lf %lf", step, cnt.n > 0 ? E / cnt.n : 0., E);
	if(u.npt) printf(" %lf", cnt.n / cnt.volume);
	printf("\n");

	pp_destroy(u.ctx);
	if(u.net) pp_net_destroy(u.net);
	free(u.r);
	free(u.O);
	free(u.bcount);
	free(u.bpart);
	ppo_free(&u.out);
	ppi_free(in);
	return 0;
}
