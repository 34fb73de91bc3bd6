/*
 * pp.h -- C ABI of libpp_hip.so, the MI355X (gfx950) implementation of the
 * PatchyParticles rototranslation Monte Carlo hot path: the Kern-Frenkel pair
 * test, the per-particle energy with cell-list neighbour search and the cell
 * rebuild, driven as checkerboard sweeps on the GPU.
 *
 * Plain C types only: pointers, sizes, doubles.  Host arrays are always in
 * particle-index order, positions as r[3*i+k], orientation matrices by rows
 * as O[9*i+3*row+col] (row k of the reference's PatchyParticle::orientation,
 * defs.h:47-56).
 *
 * Every call is synchronous with respect to the host and returns a pp_status;
 * pp_last_error() gives a human-readable message for the last failure on the
 * calling thread.  A context is not thread-safe: one thread at a time.
 *
 * Which reference interface each entry point replaces (file:line in the
 * reference, PatchyParticles):
 *   pp_create / pp_create_batch   system_init + cells_init/cells_fill + MC_init
 *                                 (system.c:58-251, cells.c:17-63, MC.c:85-131)
 *   pp_sweeps                     syst.do_ensemble called once per step
 *                                 (main.c:119 -> do_NVT MC.c:19-29 /
 *                                 do_NPT MC.c:60-68, MC_move_rototranslate
 *                                 MC.c:311-340, MC_change_volume MC.c:450-513)
 *   pp_energy                     sum over particles of MC_energy
 *                                 (MC.c:272-309; main.c:54-61)
 *   pp_check_energy               MC_check_energy + cells_check
 *                                 (MC.c:766-784, cells.c:65-76; main.c:86-87)
 *   pp_download / pp_upload(_n)   the System particle arrays read by
 *                                 output_save (output.c:302-317) / written by
 *                                 the .rrr loader (system.c:213-245)
 *   pp_snapshot_start / _finish   output_save's read of the particle arrays,
 *                                 overlapped with the following sweeps
 *   pp_get_counters(_all)         syst->tries/accepted/energy/V/box read by
 *                                 output_print (output.c:119-196)
 *   pp_reset_counters             utils_reset_acceptance_counters (utils.c:276-282)
 *   pp_energy_snapshot            MC_energy over every particle of a given
 *                                 snapshot, without a context (MC.c:272-309)
 *   pp_generate                   the generator program's random insertion
 *                                 (generator.c:10-124)
 *   pp_autotune, pp_set_trace/pp_get_trace, pp_set_timing/pp_get_timing,
 *   pp_get_grid, pp_get_layout    no reference counterpart (launch-shape
 *                                 tuning, per-sweep energy series, kernel
 *                                 timing, grid introspection)
 *   pp_colour_conflicts, pp_colour_check
 *                                 no reference counterpart (the reference is
 *                                 serial): the checkerboard invariant of the
 *                                 concurrent colour passes, as a debug check
 *   pp_bond_list                  output_print_bonds (output.c:213-268)
 *   pp_set_lx_move                the Lx_move keys and MC_change_Lx
 *                                 (system.c:170-196, MC.c:515-585)
 *   pp_set_lx_acceptance          MC_change_Lx's acceptance test (MC.c:562),
 *                                 or Metropolis in its place
 *   pp_checkpoint_save/_load     .rrr restart (output.c:37-46, system.c:213-245),
 *                                 exact and with the random-number counter
 *   pp_net_*, pp_get_learn_counters
 *                                 no reference counterpart: the learned
 *                                 proposal of the rototranslation move, a
 *                                 network-driven alternative to the uniform
 *                                 displacement/rotation of MC_move_rototranslate
 *                                 (MC.c:138-161) inside the same Metropolis test
 */
#ifndef PP_H
#define PP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pp_status {
	PP_OK = 0,
	PP_ERR_ARG = 1,         /* invalid argument */
	PP_ERR_HIP = 2,         /* HIP runtime error (no device, launch failure...) */
	PP_ERR_OVERLAP = 3,     /* configuration contains an overlap (main.c:59, MC.c:779) */
	PP_ERR_BOX = 4,         /* box too small for the cell grid (< 4 (1+delta) per side) */
	PP_ERR_CELL_CAP = 5,    /* a sweep cell holds more particles than a team stages (24) */
	PP_ERR_ENERGY = 6,      /* tracked energy disagrees with the recomputed one (MC.c:780-782) */
	PP_ERR_NOMEM = 7,       /* host or device allocation failed */
	PP_ERR_UNSUPPORTED = 8, /* ensemble other than 0 (NVT) / 5 (NPT), or Dynamics != 0 */
	PP_ERR_CELLS = 9        /* cell population != N (cells.c:74) */
} pp_status;

/* Ensembles (defs.h:21-25). */
#define PP_NVT 0
#define PP_GC 1
#define PP_NPT 5

/* Physical and move parameters of one system, the reference's input keys. */
typedef struct pp_params {
	int ensemble;              /* Ensemble: PP_NVT, PP_GC or PP_NPT */
	double temperature;        /* Temperature */
	double disp_max;           /* Disp_max: displacement uniform in +-Disp_max/2 per axis */
	double theta_max;          /* Theta_max: rotation angle uniform in [0, Theta_max) */
	double kf_delta;           /* KF_delta: bond range 1 + delta */
	double kf_cosmax;          /* KF_cosmax: patch half-angle cosine */
	double pressure;           /* P (NPT) */
	double rescale_factor_max; /* rescale_factor_max (NPT): ln V step */
	double activity;           /* Activity z (GC, system.c:91) */
	int gc_n_max;              /* GC_N_max (GC, system.c:94): most particles the replica may hold */
} pp_params;

/* Options of the GPU engine (no reference counterpart). */
typedef struct pp_options {
	uint64_t seed;           /* Philox key; the whole trajectory is a function of it */
	int threads_per_cell;    /* lanes that work on one sweep cell: 16, 32 or 64 (0: 32); 16 needs
	                            at most 16 particles per sweep cell and the per-pass launch form */
	int block_size;          /* threads per workgroup: 64/128/256 (0: 256) */
	double cell_occupancy;   /* target particles per sweep cell; 0: smallest cells (side >= 1+delta) */
	int device;              /* HIP device ordinal */
	int launch;              /* PP_LAUNCH_*: how the 8 colour passes of a sweep are launched */
	int trials_per_cell;     /* 0: each sweep cell makes as many trials as it holds, so a sweep is
	                          * exactly N attempts (MC.c:21); s in [1, 16]: each non-empty cell
	                          * makes s trials per pass, and a sweep s x (non-empty cells) attempts
	                          * (and s NPT volume attempts instead of one) */
} pp_options;

/* pp_options.launch.  On MI355X a grid barrier across the 8 XCDs costs more
 * than a kernel boundary, so the default is one launch per colour pass. */
#define PP_LAUNCH_DEFAULT 0   /* PP_LAUNCH_PER_PASS */
#define PP_LAUNCH_PER_PASS 1  /* one launch per colour pass */
#define PP_LAUNCH_COOP 2      /* one cooperative launch per sweep (grid barriers between passes); fail if unsupported */
#define PP_LAUNCH_DEPS 3      /* one launch per sweep in which each cell starts its pass as soon as its neighbours
                                 of earlier passes are done; 32 or 64 threads per cell.  Results equal the
                                 per-pass form's bit for bit (learned passes still launch per pass) */

typedef struct pp_counters {
	long long tries[2];    /* [0] rototranslations, [1] volume moves, since the last reset */
	long long accepted[2];
	long long bonds;       /* tracked number of bonded pairs: energy = -bonds */
	long long sweeps;      /* sweeps done since creation */
	double box[3];
	double volume;         /* V as the reference tracks it (MC.c:500) */
	long long gc_tries[2];    /* GC: [0] insertions, [1] removals (ADD, REMOVE of defs.h:13-14) */
	long long gc_accepted[2];
	int n;                 /* particles in the replica now (GC: changes; otherwise fixed) */
	int pad;
} pp_counters;

typedef struct pp_ctx pp_ctx;

/* One system.  r: 3n doubles, O: 9n doubles (rows), box: 3 doubles.
 * Fails with PP_ERR_OVERLAP if the configuration overlaps (main.c:57-60). */
pp_status pp_create(const pp_params *params, const pp_options *opts, int n, const double *r,
                    const double *O, const double box[3], pp_ctx **out);

/* n_rep independent systems in one context; all kernels run on all of them
 * at once.  params[n_rep], n[n_rep]; r, O concatenated in replica order;
 * box[3*n_rep]. */
pp_status pp_create_batch(int n_rep, const pp_params *params, const pp_options *opts, const int *n,
                          const double *r, const double *O, const double *box, pp_ctx **out);

/* Advance every replica by n sweeps.  A sweep is N single-particle
 * rototranslation attempts (MC.c:21) done as 8 checkerboard colour passes,
 * plus, under NPT, one volume-move attempt (expected once per sweep in
 * do_NPT, MC.c:62-63), and, for a grand-canonical replica (PP_GC), one
 * phase of insertions and removals (MC_add_remove, MC.c:342-448) in which
 * every cell of one random colour tries one of the two, accepted with the
 * reference's factors for the cell's volume and population:
 * z v_c / (n_c + 1) exp(-dE/T) and n_c / (z v_c) exp(-dE/T).  That phase
 * needs n_c free slots beyond the particles (GC_N_max >= N + cells of one
 * colour), else the sweep fails with PP_ERR_ARG. */
pp_status pp_sweeps(pp_ctx *ctx, long long n);

/* Recompute the energy of replica rep from scratch: bonds[i] = number of
 * bonded neighbours of particle i, overlap[i] = 1 if i overlaps a neighbour
 * (then bonds[i] = 0, as MC_energy returns 0).  Either array may be NULL.
 * total_bonds = number of bonded pairs, any_overlap = OR of overlap[]. */
pp_status pp_energy(pp_ctx *ctx, int rep, int32_t *bonds, uint8_t *overlap, long long *total_bonds,
                    int *any_overlap);

/* MC_check_energy for every replica: recompute, compare with the tracked
 * energy (relative tolerance 1e-5), re-orthonormalise every orientation by
 * Gram-Schmidt, count the cell population.  energy_per_rep may be NULL. */
pp_status pp_check_energy(pp_ctx *ctx, double *energy_per_rep);

/* Particles of replica rep in index order (a grand-canonical replica: its
 * pp_counters.n particles, renumbered 0 .. n - 1 in the order of their
 * indices; removals and insertions change the numbering as the reference's
 * removals do, MC.c:405-433). */
pp_status pp_download(pp_ctx *ctx, int rep, double *r, double *O, double box[3]);
pp_status pp_upload(pp_ctx *ctx, int rep, const double *r, const double *O, const double box[3]);
/* pp_upload with the particle count given: a grand-canonical replica takes
 * any count up to its GC_N_max, any other replica only its own. */
pp_status pp_upload_n(pp_ctx *ctx, int rep, int n, const double *r, const double *O, const double box[3]);

pp_status pp_get_counters(pp_ctx *ctx, int rep, pp_counters *out);
/* pp_get_counters for every replica at once: out[pp_num_replicas(ctx)]. */
pp_status pp_get_counters_all(pp_ctx *ctx, pp_counters *out);
pp_status pp_reset_counters(pp_ctx *ctx);
int pp_num_replicas(const pp_ctx *ctx);
/* Particles of replica rep; for a grand-canonical replica the most it may
 * hold (GC_N_max): its current count is pp_counters.n. */
int pp_num_particles(const pp_ctx *ctx, int rep);

/* Per-particle energy of a snapshot without a context (GPU, same kernels). */
pp_status pp_energy_snapshot(int n, const double *r, const double *O, const double box[3], double kf_delta,
                             double kf_cosmax, int32_t *bonds, uint8_t *overlap, long long *total_bonds,
                             int *any_overlap);

/* pp_energy_snapshot without allocating: no counterpart in the reference;
 * for callers that cannot allocate (code under stream capture, serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc or
 * hipFree, and allocates no host memory except the error message of a
 * failed call.  Same arguments, results and device (0) as
 * pp_energy_snapshot, for 0 <= n <= PP_NOALLOC_MAX_N (PP_ERR_ARG beyond).
 * Its device memory is one fixed array of about 16 MB per GPU, held until
 * the process ends and shared by every call on that GPU: calls that overlap
 * on one GPU race. */
#define PP_NOALLOC_MAX_N 65536
pp_status pp_energy_snapshot_noalloc(int n, const double *r, const double *O, const double box[3], double kf_delta,
                                     double kf_cosmax, int32_t *bonds, uint8_t *overlap, long long *total_bonds,
                                     int *any_overlap);

/* Random sequential insertion of n non-overlapping unit spheres in a cubic
 * box of side (n/rho)^(1/3), each with a uniformly random orientation
 * (generator.c:43-84, the reference's generator program; host code, no GPU).
 * seed < 0 reproduces the reference generator's stream; rho <= 0.7.
 * r: 3n, O: 9n, box: 3 doubles. */
pp_status pp_generate(int n, double rho, long seed, double *r, double *O, double box[3]);

/* Energy trace: with capacity > 0, the bonded-pair count of every replica is
 * recorded after each of the next `capacity` sweeps (the energy.dat series
 * of output_print, output.c:124, at every sweep); capacity 0 turns it off.
 * pp_get_trace copies up to `capacity` entries, oldest first. */
pp_status pp_set_trace(pp_ctx *ctx, int capacity);
pp_status pp_get_trace(pp_ctx *ctx, int rep, long long *bonds, int capacity, int *count);

/* Kernel timing with HIP events on the context's stream (off by default).
 * every = 0: off; every = k >= 1: the launches of every k-th sweep (and every
 * energy call) are bracketed by events.  Resets the accumulated times.
 * name: "pass", "rebuild", "energy", "volume", "learned" or NULL for the sum. */
pp_status pp_set_timing(pp_ctx *ctx, int every);
pp_status pp_get_timing(pp_ctx *ctx, const char *name, long long *launches, double *total_ms);

/* Autotuner (no reference counterpart): runs `sweeps` sweeps (after one
 * warm-up sweep) with each feasible launch shape -- threads per cell 32 / 64
 * times workgroups of 64 / 128 / 256 threads -- and keeps the fastest for the
 * context.  The sweeps are real: the trajectory does not depend on the launch
 * shape, so tuning only advances the simulation.  Outputs may be NULL.
 * Per-pass launch form only. */
pp_status pp_autotune(pp_ctx *ctx, int sweeps, int *threads_per_cell, int *block_size, double *us_per_sweep);

/* Box-shape moves (MC_change_Lx, MC.c:515-585; keys Lx_move, Lx_change_max,
 * Lyz_min, Lyz_max, system.c:170-196) for replica rep (-1: all), NVT only:
 * about once per N particle attempts, Lx += U(-1/2, 1/2) lx_change_max with
 * Ly = Lz = sqrt(V / Lx), rejected outside [lyz_min, lyz_max], accepted with
 * min(1, exp(-dE/T)).  Their tries / acceptances are counted in tries[1] /
 * accepted[1] of pp_counters.  lx_change_max <= 0 turns them off.  Errors
 * as the reference's: Ly != Lz, lyz_min >= Ly, lyz_max <= Ly. */
pp_status pp_set_lx_move(pp_ctx *ctx, int rep, double lx_change_max, double lyz_min, double lyz_max);

/* The acceptance rule of the box-shape moves of replica rep (-1: all).
 * PP_LX_METROPOLIS (the default): min(1, exp(-dE/T)), which keeps the
 * Boltzmann distribution.  PP_LX_REFERENCE: the reference's own test,
 * u < exp(+dE/T) with dE = E_new - E_old (MC_change_Lx, MC.c:562), for runs
 * that must reproduce the reference program's box-shape statistics; it
 * favours moves that raise the energy. */
#define PP_LX_METROPOLIS 0
#define PP_LX_REFERENCE 1
pp_status pp_set_lx_acceptance(pp_ctx *ctx, int rep, int rule);

/* Bonded partners of every particle of replica rep, the pairs of
 * output_print_bonds (output.c:213-268, MC_interact's test):
 * count[i] = number of particles bonded to particle i, and their indices at
 * partners[i * max_per + k] for k < min(count[i], max_per), in no particular
 * order (the host program sorts them into the reference's cell order). */
pp_status pp_bond_list(pp_ctx *ctx, int rep, int max_per, int32_t *count, int32_t *partners);

/* Checkpoint (no reference counterpart; the reference restarts from .rrr
 * text, output.c:37-46 and system.c:213-245, which rounds the state): the
 * exact device state of every replica (FP64 positions and orientations, box,
 * tracked volume, move counters) and the sweep index, which is the Philox
 * counter.  Loading into a context created with the same parameters, sizes,
 * seed and trials_per_cell resumes the run bit for bit. */
pp_status pp_checkpoint_save(pp_ctx *ctx, const char *path);
pp_status pp_checkpoint_load(pp_ctx *ctx, const char *path);

/* Asynchronous snapshot of replica rep: the state output_save writes
 * (output.c:302-317, main.c:98-125), taken without holding the sweeps back.
 * pp_snapshot_start copies the replica's particles on the device into a
 * staging area of the context and starts their copy to pinned host memory on
 * a second stream, then returns; sweeps enqueued after it run while the copy
 * proceeds.  pp_snapshot_finish waits for the copy and writes r (3n), O (9n)
 * and box as pp_download does, in particle-index order, for the state at the
 * time of the start.  One snapshot per context at a time: a second start
 * before the finish is refused. */
pp_status pp_snapshot_start(pp_ctx *ctx, int rep);
pp_status pp_snapshot_finish(pp_ctx *ctx, double *r, double *O, double box[3]);

/* Grid of replica rep: cells per side and cell side lengths. */
pp_status pp_get_grid(pp_ctx *ctx, int rep, int n_side[3], double side[3]);
/* Cell layout of the context: the cells reserved over all replicas (NPT
 * replicas reserve room for their grids to grow), and the tiles the
 * rebuild's scan takes them in (scan_tile cells each, scan_tiles of them). */
pp_status pp_get_layout(pp_ctx *ctx, long long *cell_capacity, int *scan_tile, int *scan_tiles);

/* The checkerboard invariant (a debug check).  The particles of one colour's
 * cells move concurrently, so none may lie within the bond range 1 + delta of
 * a particle in another cell of the same colour (colour = the parities of the
 * cell's three indices).  pp_colour_conflicts counts the particles of replica
 * rep that do, on the grid of its current cell lists (nside NULL) or on
 * nside[0] x nside[1] x nside[2] cells over its box without shift (any grid:
 * a grid with an odd count or with cells narrower than 1 + delta has
 * conflicts); brute force over the replica's pairs, minimum image. */
pp_status pp_colour_conflicts(pp_ctx *ctx, int rep, const int nside[3], long long *count);
/* Live checks: with on = 1 every later sweep counts the conflicts of every
 * replica on its lists' grid after the rebuild and after each colour pass
 * (after the whole launch in the PP_LAUNCH_COOP and PP_LAUNCH_DEPS forms), and
 * adds them to a total (reset here); on = 0 stops the checks, on = -1 only
 * reads.  *total and *checks (each may be NULL) receive the conflicts found
 * and the replica checks made since the last reset. */
pp_status pp_colour_check(pp_ctx *ctx, int on, long long *total, long long *checks);

/* ---- learned proposal (no reference counterpart) --------------------------
 * A move of particle p: the descriptor of its neighbourhood (the 12 nearest
 * particles within 1.5 among the 27 cells around p's sweep cell: separation in
 * p's body frame and the quaternion of the relative orientation) -> an MLP
 * 84 (padded to 96) -> 128 -> 128 -> 128 -> 26 (ReLU; BF16 weights, FP32
 * accumulation) -> a Gaussian displacement in p's body frame and a
 * 4-component von Mises-Fisher mixture over the rotation axis; the angle is
 * uniform in [0, theta_max).  The move is accepted with
 * min(1, exp(-dE/T) q(x'->x) / q(x->x')), so the stationary distribution is
 * the Boltzmann one of the Kern-Frenkel model, as with the plain move.
 *
 * Weights: layer l of W is an nn_k(l) x nn_n(l) row-major block (inputs x
 * outputs; k = 96, 128, 128, 128; n = 128, 128, 128, 32), the blocks one after
 * another (PP_NET_WEIGHTS floats); biases likewise (PP_NET_BIASES floats).
 * Rows of padded inputs (84..95) and columns of padded outputs (26..31) are
 * ignored by the forward pass only if kept zero; pp_net_create zeroes them.
 * Output layout: mu[3], log sigma[3], then per component c < 4: direction[3]
 * (normalised by the density), log kappa, logit. */
#define PP_NET_INPUTS 84
#define PP_NET_OUTPUTS 26
#define PP_NET_WEIGHTS 49152
#define PP_NET_BIASES 416

typedef struct pp_net pp_net;

/* A network on `device` with Glorot-uniform weights from `seed`, its last
 * layer scaled so that the untrained proposal is close to a small isotropic
 * move (sigma 0.03). */
pp_status pp_net_create(int device, uint64_t seed, pp_net **out);
pp_status pp_net_set(pp_net *net, const float *w, const float *b);
pp_status pp_net_get(pp_net *net, float *w, float *b);
/* Forward pass of m descriptors x[m][PP_NET_INPUTS] (rounded to BF16 as the
 * sampler does) into y[m][PP_NET_OUTPUTS]. */
pp_status pp_net_forward(pp_net *net, int m, const float *x, float *y);
/* Mean over m rows of -log q(target | x), target[m][6] = body-frame
 * displacement[3] and unit rotation axis[3], and its gradient with respect to
 * the weights (gw: PP_NET_WEIGHTS, gb: PP_NET_BIASES; outputs may be NULL). */
pp_status pp_net_loss_grad(pp_net *net, int m, const float *x, const float *target, double *loss, float *gw,
                           float *gb);
void pp_net_destroy(pp_net *net);
/* Descriptors (as the learned move computes them, BF16 values) of particles
 * idx[0..m) of a configuration of n particles in `box`, every other particle a
 * candidate neighbour: x[m][PP_NET_INPUTS]. */
pp_status pp_net_descriptor(int device, int n, const double *r, const double *O, const double box[3], int m,
                            const int *idx, float *x);

/* Use `net` for a share `learned_fraction` of the colour passes of every
 * later sweep (chosen per pass from the seed, the sweep and the pass index;
 * the other passes are the plain ones).  A learned pass moves one member of
 * each non-empty cell.  net NULL detaches.  The network must outlive the
 * attachment. */
pp_status pp_net_attach(pp_ctx *ctx, pp_net *net, double learned_fraction, double theta_max);
/* Train `net` on the context's replicas: each step draws `batch` particles,
 * perturbs each (Gaussian displacement of deviation pos_noise, rotation by an
 * angle uniform in [0, theta_max) about a uniform axis) and takes one Adam
 * step (rate lr) on the mean -log q of the move back; `sweeps_between` sweeps
 * run between steps.  Outputs the mean loss of the first and the last batch
 * (before their update); outputs may be NULL. */
pp_status pp_net_train(pp_ctx *ctx, pp_net *net, int steps, int batch, double lr, double pos_noise,
                       double theta_max, int sweeps_between, double *loss_first, double *loss_last);
/* Learned moves tried / accepted by replica rep since the last reset. */
pp_status pp_get_learn_counters(pp_ctx *ctx, int rep, long long *tries, long long *accepted);

void pp_destroy(pp_ctx *ctx);
const char *pp_strerror(pp_status s);
const char *pp_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
