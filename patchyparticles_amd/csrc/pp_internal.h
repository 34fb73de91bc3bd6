// pp_internal.h -- structures shared by the gfx950 kernels (pp_kernels.hip)
// and the C-ABI host code (pp_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "kf_device.hpp"

namespace ppd {

constexpr int CELL_CAP = 24;   // particles of one sweep cell staged in LDS
constexpr int NSLOT = 64;      // spread of per-replica atomic counters
constexpr int NCOMP = 12;      // x, y, z, O00..O22 per particle
constexpr int TRIAL_W = 14;    // one trial: member rank, dx, dy, dz, R[9], acceptance draw
constexpr int MAX_TRIALS_PER_CELL = 16;
// cells per scan workgroup of 256 lanes: small tiles shorten each lane's
// chain where the whole grid is a few tiles; large ones keep the per-tile
// overheads down on grids of many tiles
constexpr int SCAN_TILE_SMALL = 512;
constexpr int SCAN_TILE_LARGE = 2048;
constexpr int SCAN_SMALL_MAX_TILES = 320;  // small tiles up to this many (look-back scan)
constexpr int REC_W = 64;      // ints of one cell record of a pass
constexpr int ACC_TAB = 16;    // Metropolis factors exp(-k/T) tabulated per replica, k < ACC_TAB

// per-replica counter rows, each NSLOT wide
enum Ctr : int {
	C_TRY_ROTO = 0, C_ACC_ROTO, C_DBONDS, C_TRY_VOL, C_ACC_VOL, C_EBONDS, C_OVL, C_TRY_LEARN, C_ACC_LEARN,
	C_TRY_ADD, C_ACC_ADD, C_TRY_REM, C_ACC_REM, NCTR
};

// ------------------------------------------------------------------------
// learned proposal network: descriptor of up to NN_NB neighbours within
// NN_RCUT (body-frame position, relative orientation quaternion), an MLP
// NN_IN_PAD -> 128 -> 128 -> 128 -> NN_OUT_PAD (ReLU), BF16 weights, FP32
// accumulation, and the outputs of a Gaussian displacement and a 4-component
// von Mises-Fisher mixture over the rotation axis.
constexpr int NN_NB = 12;
constexpr int NN_FEAT = 7;
constexpr int NN_IN = NN_NB * NN_FEAT;  // 84
constexpr int NN_IN_PAD = 96;
constexpr int NN_H = 128;
constexpr int NN_OUT = 26;  // mu[3], log sigma[3], 4 x (direction[3], log kappa, logit)
constexpr int NN_OUT_PAD = 32;
constexpr int NN_MIX = 4;
constexpr double NN_RCUT = 1.5;
constexpr int NN_LAYERS = 4;
__host__ __device__ constexpr int nn_k(int l) { return l == 0 ? NN_IN_PAD : NN_H; }
__host__ __device__ constexpr int nn_n(int l) { return l == NN_LAYERS - 1 ? NN_OUT_PAD : NN_H; }
__host__ __device__ constexpr int nn_woff(int l) { return l == 0 ? 0 : nn_woff(l - 1) + nn_k(l - 1) * nn_n(l - 1); }
__host__ __device__ constexpr int nn_boff(int l) { return l == 0 ? 0 : nn_boff(l - 1) + nn_n(l - 1); }
constexpr int NN_W = nn_woff(NN_LAYERS);  // 49152 weights
constexpr int NN_B = nn_boff(NN_LAYERS);  // 416 biases
// clamps of the proposal parameters (the density is that of the clamped values)
constexpr float NN_LS_MIN = -6.2f;   // log sigma: 0.002 .. 0.6
constexpr float NN_LS_MAX = -0.5f;
constexpr float NN_LK_MIN = -3.0f;   // log kappa: 0.05 .. 3000
constexpr float NN_LK_MAX = 8.0f;

// Device state of one network: FP32 master weights and Adam moments, and the
// BF16 operand fragments of the MFMA kernels (forward: W; backward: W^T).
struct NetDev {
	float *w, *b;
	float *mw, *vw, *mb, *vb;
	float *gw, *gb;
	__bf16 *frag, *fragT;
};

constexpr int NN_PROP_W = 16;  // doubles of one drawn proposal

// Per-context state of the learned passes (scratch sized for the largest list).
struct LearnDev {
	int on;
	double theta_max;    // rotation angle of a learned move, uniform in [0, theta_max)
	__bf16 *xf, *xr;     // [total_act][NN_IN_PAD] descriptors at the current / proposed state
	float *yf, *yr;      // [total_act][NN_OUT_PAD] network outputs
	double *prop;        // [total_act][NN_PROP_W] p'[3], O'[9], log q forward, a[3] of the proposal
	int *slot;           // [total_act] moving slot (-1: no move)
};

// device error bits
enum Err : int { ERR_CELL_CAP = 1, ERR_BOX = 2, ERR_SCAN = 8, ERR_BARRIER = 16, ERR_GC_FULL = 32, ERR_DEPWAIT = 64 };

// rebuild modes
enum Reb : int { REB_KEEP = 0, REB_SHIFT = 1, REB_TRIAL = 2 };

// a decoupled look-back serves at most this many tiles well (its chains grow
// with the tiles in flight); grids of large tiles, which always have more,
// are scanned in three steps (tile sums, one scan of the sums, the tiles)
constexpr int SCAN_LOOKBACK_MAX_TILES = 64;

// spin bound of the scan's look-back (a stuck predecessor is reported, not waited on forever)
constexpr int SCAN_SPIN_MAX = 1 << 22;
// polls of a grid barrier before it reports ERR_BARRIER and lets the workgroup go on
constexpr int BAR_SPIN_MAX = 1 << 21;
// polls of a dependency-driven sweep's wait for neighbour cells before it
// reports ERR_DEPWAIT and lets the team go on
constexpr int DD_SPIN_MAX = 1 << 20;

// A cell grid over one replica's (periodic, orthorhombic) box.
struct Grid {
	double box[3];
	double inv_box[3];
	double shift[3];
	int n[3];
	int pad;
};

// One replica: its particle range, its cell region and its parameters.
//
// Slots [p_off, p_off + cap_np) belong to the replica; after a rebuild its np
// particles fill [p_off, p_off + np) grouped by cell, and the rest of its
// slots sit in its parking cell (local index cell_cap, after the grid's
// cells, never listed or searched).  Only grand-canonical replicas have
// cap_np > np: their insertions take free slots (np_hw: slots in use since
// the last rebuild) and their removals mark slots dead (particle index -1).
struct Rep {
	Grid g;   // grid of the current cell structure
	Grid tg;  // grid of an NPT trial (valid while trial_active)
	int p_off, np;
	int cap_np, np_hw;
	int gc;           // grand-canonical insertions and removals (Ensemble = 1)
	int gc_base;      // np when this sweep's GC phase began (-1: the phase is skipped)
	int gc_colour;    // colour of this sweep's GC phase
	int pad3_;
	double activity;  // z (system.c:91)
	int cell_off, cell_cap;
	int cap_n[3];
	int act_cap;       // cap_n[0]/2 * cap_n[1]/2 * cap_n[2]/2: cells of one colour at most
	int sel;           // buffer that holds the current state (0/1)
	int ensemble;
	int min_n;         // smallest cells per side of the grid (4 for sweeps)
	int even;          // sweep grids need an even count per side
	int trial_active;  // an NPT trial is in flight
	int trial_dir;
	uint32_t col_perm;  // colour of pass k in bits 4k..4k+3 (this sweep)
	int pad_;
	double side_min;   // smallest allowed cell side
	double T, disp_max, theta_max, rc2, cosmax, P, rfm;
	double V;          // the volume the reference tracks (MC.c:500)
	long long bonds_base;  // bonded pairs = bonds_base + sum of C_DBONDS slots
	double trial_f, trial_V, trial_V0;
	double trial_s[3];        // scale of each position component in the trial (1: unchanged)
	int trial_kind;           // 0: volume move (NPT), 1: Lx move (NVT with box-shape moves)
	int lx_move;              // box-shape moves on (MC_change_Lx, MC.c:515-585)
	// The grid the cell lists of the current buffer were built on, in the
	// current coordinates: membership by lg.  It differs from g after a box
	// move accepted without a rebuild (the lists keep their members, lg is
	// rescaled with the positions).  lists_ok = 0: the lists describe no
	// grouping of the current buffer (a rejected trial rebuilt them into the
	// spare one) until the next rebuild.
	Grid lg;
	int lists_ok;
	int trial_fast;  // the trial in flight is evaluated on the current lists (no trial rebuild)
	// rec_ok: the cell records of this sweep's passes describe the current
	// lists (cleared by any other rebuild); a fast trial then takes its
	// energy from the records, pair by pair (trial_rec)
	int rec_ok;
	int trial_rec;
	// pl_ok: the sweep's list of close pairs (built from the records at the
	// start of the box trials) holds this replica's pairs; pl_scale: the
	// product of the axis scalings accepted since; a fast trial whose
	// smallest cumulative scaling keeps every pair beyond the list radius
	// outside the bond range takes its energy from that list (trial_pl)
	int pl_ok;
	int trial_pl;
	double pl_scale[3];
	int vol_apply;   // an accepted fast trial: positions still to be rescaled in place
	int lx_ref_rule;  // box-shape moves accepted with exp(+dE/T) as MC.c:562 (not Metropolis)
	double lx_change_max, lyz_min, lyz_max;
	double acc_tab[ACC_TAB];  // exp(-k / T): the acceptance factor of an energy rise of k bonds
};

// Cells per side for a box: the largest (even, if asked) count whose side is
// at least side_min, capped per axis.  Returns false if below min_n.
__host__ __device__ inline bool make_grid(const double box[3], double side_min, int min_n, int even,
                                          const int cap[3], int n[3]) {
	bool ok = true;
	for(int k = 0; k < 3; k++) {
		int m;
		if(even) m = 2 * (int) floor(box[k] / (2. * side_min));
		else m = (int) floor(box[k] / side_min);
		if(cap && m > cap[k]) m = even ? (cap[k] & ~1) : cap[k];
		if(m < min_n) ok = false;
		n[k] = m;
	}
	return ok;
}

__host__ __device__ inline void set_grid_box(Grid &G, const double box[3]) {
	for(int k = 0; k < 3; k++) {
		G.box[k] = box[k];
		G.inv_box[k] = 1. / box[k];
	}
}

// Device buffers of a context.
struct DevState {
	double *f[2];        // SoA particle data, component c of slot j at f[b][c*stride + j]
	int *pid[2];         // particle index of slot j
	int stride;
	int Ntot;
	int R;
	Rep *reps;
	const int *prep;     // replica of slot j (fixed)
	const int *act_off;  // prefix over replicas of active-cell capacity per colour (R+1)
	int total_act;
	int C_total;         // total cell capacity
	int *count, *start, *slot, *cellid;
	unsigned long long *scan_status;  // per tile: flag << 32 | value (decoupled look-back)
	int *clist;        // non-empty cells by colour: replica r, colour k at cell_off + k * act_cap
	int *ccount;       // [R][8] non-empty cells per colour
	int *pfx;          // [8][R+1] prefix over replicas of the cells pass k works on
	int *rec;          // [8][total_act][REC_W] records of the cells of each pass, in list order
	unsigned *bar;     // grid barrier of the cooperative sweep kernel: arrivals, generation
	int *scan_ticket;  // [0]: tiles taken, [1]: tiles done
	int *tile_sum;     // [ntiles] per-tile sums, then their exclusive prefixes (grids of many tiles)
	int ntiles;
	int scan_tile;     // SCAN_TILE_SMALL or SCAN_TILE_LARGE
	unsigned long long *ctr;  // [R][NCTR][NSLOT]
	int *err;
	int32_t *out_bonds;
	uint8_t *out_ovl;
	int trials_per_cell;  // trials of a non-empty cell in one pass; 0: as many as it holds
	LearnDev learn;
	long long *trace;  // [R][trace_cap] bonded pairs at the start of each sweep (optional)
	int trace_cap;
	uint64_t seed;
	int *vflag;        // [R] a trial energy found an overlap: the replica's other teams stop
	int *vslow;        // [2] the trial paths the attempt's replicas need (TrialNeed bits); the attempt
	                   // at which a run-ahead sequence of attempts stopped, -1 while none did
	// how box trials find their energy: 0 the fastest path available (the
	// close-pair list, else the cell records, else the current lists per
	// particle, else a trial rebuild); 3 never the pair list; 1 neither the
	// pair list nor the records; 2 always a trial rebuild (diagnostic forms
	// that must give the same trajectories)
	int npt_trial_form;
	unsigned *done;    // [C_total] dependency-driven sweeps: mark of the sweep whose pass the cell last finished
	int *ddq;          // [1] dependency-driven sweeps: tickets taken (reset by the rebuild)
	unsigned long long *ccheck;  // [1] checkerboard conflicts found by the live checks
	int *ovf;          // [total_act] list entries of a pass deferred by the small-cell pass to the general one
	int *ovf_n;        // [8] entries deferred per pass of the sweep (reset by the rebuild)
	int small_cap;     // members from which the small-cell pass defers a cell (at most its LDS capacity)
	int2 *pl;          // [pl_cap] the close pairs (slot j, slot q) of the box-trial replicas, this sweep
	int *pl_n;         // [2] pairs listed, and 1 if the list overflowed
	int pl_cap;
};

// launchers (pp_kernels.hip)
int small_cell_cap();  // members a small-cell pass team stages in LDS
void launch_fill_i32(int *p, int v, long long n, hipStream_t s);
void launch_fill_u64(unsigned long long *p, unsigned long long v, long long n, hipStream_t s);
// regroup slots by cell (mode REB_*); REB_SHIFT draws the grid shift of
// `sweep` and records the energy trace entry trace_pos (if >= 0)
void launch_rebuild(const DevState &d, int mode, uint64_t sweep, long long trace_pos, hipStream_t s);
void launch_pass(const DevState &d, uint64_t sweep, int pass, int T, int block, hipStream_t s);
// the 8 colour passes of a sweep in one dependency-driven launch; cells
// publish `mark` (never 0, different from the previous sweep's) when done
void launch_passes_dd(const DevState &d, uint64_t sweep, unsigned mark, int T, int block, hipStream_t s);
// learned proposal (pp_learn.hip)
void launch_learned_pass(const DevState &d, const NetDev &net, uint64_t sweep, int pass, hipStream_t s);
void launch_net_forward(const NetDev &net, const __bf16 *x, int m, const int *m_dev, float *y, __bf16 *h,
                        hipStream_t s);
void launch_net_pack(const NetDev &net, hipStream_t s);
// m rows (padded to a multiple of 32, rows >= m of x zero): forward with the
// hidden layers kept, per-row -log q (loss[m_pad]), backward, gradients of
// the mean loss into net.gw / net.gb
void launch_net_loss_grad(const NetDev &net, const __bf16 *x, const float *target, int m, __bf16 *h, float *y,
                          __bf16 *dy, __bf16 *dh, float *loss, hipStream_t s);
void launch_net_adam(const NetDev &net, int step, float lr, hipStream_t s);
void launch_descriptor(int n, const double *r, const double *O, const Grid &g, int m, const int *idx, __bf16 *x,
                       hipStream_t s);
void launch_train_samples(const DevState &d, uint64_t seed, uint64_t step, int m, double pos_noise, double theta_max,
                          __bf16 *x, float *target, hipStream_t s);
// the 8 passes of a sweep in one cooperative launch of `blocks` workgroups
hipError_t launch_sweep_passes(const DevState &d, uint64_t sweep, int T, int block, int blocks, hipStream_t s);
int sweep_blocks_max(int T, int block, int cus);  // co-resident workgroups of that launch
size_t pass_lds_bytes(int T, int block);  // dynamic LDS of one pass workgroup
constexpr size_t MAX_LDS_BYTES = 65536;
void launch_energy(const DevState &d, int trial, int per_particle, hipStream_t s);
void launch_bond_list(const DevState &d, int max_per, int *count, int *partners, hipStream_t s);
// volume-move attempt k of sweep `sweep` (NPT replicas only)
void launch_vol_propose(const DevState &d, uint64_t sweep, int k, hipStream_t s);
// write (seq << 4) | the TrialNeed bits of the attempt just proposed into the
// host's pinned word (host_word: its device address); seq < 2^27
void launch_vol_publish(const DevState &d, int *host_word, unsigned seq, hipStream_t s);
void launch_vol_decide(const DevState &d, uint64_t sweep, int k, hipStream_t s);
// rescale the positions of replicas whose fast trial was accepted
void launch_vol_apply(const DevState &d, hipStream_t s);
// the energy of the box trials whose replicas take it from the cell records
// (trial_rec), one listed cell per team; `cus` sizes the grid
void launch_energy_rec(const DevState &d, int cus, hipStream_t s);
// checkerboard invariant (debug): add to *out the particles of replica rep
// within the bond range of a particle in another cell of the same colour, on
// the grid of its current lists (nside == nullptr) or on nside cells per axis
void launch_colour_conflicts(const DevState &d, int rep, int slots, const int *nside, unsigned long long *out,
                             hipStream_t s);
// zero the energy rows (C_EBONDS, C_OVL) of every replica
void launch_energy_rows_zero(const DevState &d, hipStream_t s);
// run-ahead box attempts: stop the sequence at attempt k when it needs more
// than the close-pair list; report the stop (attempt + 1, 0 for none) with
// `seq` into the host's pinned word and clear it
void launch_vol_gate(const DevState &d, int k, hipStream_t s);
void launch_vol_report(const DevState &d, int *host_word, unsigned seq, hipStream_t s);
// the trial-energy paths a box attempt needs, as the host reads them from
// vslow after k_vol_propose: a trial rebuild, the per-particle trial energy,
// the per-record trial energy
enum TrialNeed : int { NEED_REBUILD = 1, NEED_PARTICLES = 2, NEED_RECORDS = 4, NEED_PAIRS = 8 };
// the close-pair list keeps the pairs of adjacent cells closer than
// PAIR_LIST_FACTOR bond ranges at its build
constexpr double PAIR_LIST_FACTOR = 1.1;
// build the sweep's close-pair list (and reset the replicas' pair state)
void launch_pairs_build(const DevState &d, int cus, hipStream_t s);
// the energy of the box trials that take it from the close-pair list
void launch_energy_pairs(const DevState &d, int cus, hipStream_t s);
// grand-canonical insertions and removals of the sweep (GC replicas only)
void launch_gc_phase(const DevState &d, uint64_t sweep, hipStream_t s);
void launch_gram_schmidt(const DevState &d, hipStream_t s);

}  // namespace ppd
