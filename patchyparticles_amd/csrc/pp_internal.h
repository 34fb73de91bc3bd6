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
constexpr int SCAN_TILE = 2048;
constexpr int REC_W = 64;      // ints of one cell record of a pass
constexpr int ACC_TAB = 16;    // Metropolis factors exp(-k/T) tabulated per replica, k < ACC_TAB

// per-replica counter rows, each NSLOT wide
enum Ctr : int { C_TRY_ROTO = 0, C_ACC_ROTO, C_DBONDS, C_TRY_VOL, C_ACC_VOL, C_EBONDS, C_OVL, NCTR };

// device error bits
enum Err : int { ERR_CELL_CAP = 1, ERR_BOX = 2, ERR_SCAN = 8, ERR_BARRIER = 16 };

// rebuild modes
enum Reb : int { REB_KEEP = 0, REB_SHIFT = 1, REB_TRIAL = 2 };

// spin bound of the scan's look-back (a stuck predecessor is reported, not waited on forever)
constexpr int SCAN_SPIN_MAX = 1 << 22;
// polls of a grid barrier before it reports ERR_BARRIER and lets the workgroup go on
constexpr int BAR_SPIN_MAX = 1 << 21;

// A cell grid over one replica's (periodic, orthorhombic) box.
struct Grid {
	double box[3];
	double inv_box[3];
	double shift[3];
	int n[3];
	int pad;
};

// One replica: its particle range, its cell region and its parameters.
struct Rep {
	Grid g;   // grid of the current cell structure
	Grid tg;  // grid of an NPT trial (valid while trial_active)
	int p_off, np;
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
	int ntiles;
	unsigned long long *ctr;  // [R][NCTR][NSLOT]
	int *err;
	int32_t *out_bonds;
	uint8_t *out_ovl;
	int trials_per_cell;  // trials of a non-empty cell in one pass; 0: as many as it holds
	long long *trace;  // [R][trace_cap] bonded pairs at the start of each sweep (optional)
	int trace_cap;
	uint64_t seed;
};

// launchers (pp_kernels.hip)
void launch_fill_i32(int *p, int v, long long n, hipStream_t s);
void launch_fill_u64(unsigned long long *p, unsigned long long v, long long n, hipStream_t s);
// regroup slots by cell (mode REB_*); REB_SHIFT draws the grid shift of
// `sweep` and records the energy trace entry trace_pos (if >= 0)
void launch_rebuild(const DevState &d, int mode, uint64_t sweep, long long trace_pos, hipStream_t s);
void launch_pass(const DevState &d, uint64_t sweep, int pass, int T, int block, hipStream_t s);
// the 8 passes of a sweep in one cooperative launch of `blocks` workgroups
hipError_t launch_sweep_passes(const DevState &d, uint64_t sweep, int T, int block, int blocks, hipStream_t s);
int sweep_blocks_max(int T, int block, int cus);  // co-resident workgroups of that launch
size_t pass_lds_bytes(int T, int block);  // dynamic LDS of one pass workgroup
constexpr size_t MAX_LDS_BYTES = 65536;
void launch_energy(const DevState &d, int trial, int per_particle, hipStream_t s);
// volume-move attempt k of sweep `sweep` (NPT replicas only)
void launch_vol_propose(const DevState &d, uint64_t sweep, int k, hipStream_t s);
void launch_vol_decide(const DevState &d, uint64_t sweep, int k, hipStream_t s);
void launch_gram_schmidt(const DevState &d, hipStream_t s);

}  // namespace ppd
