// pp_internal.h -- structures shared by the gfx950 kernels (pp_kernels.hip)
// and the C-ABI host code (pp_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "kf_device.hpp"

namespace ppd {

constexpr int CELL_CAP = 32;   // particles of one sweep cell staged in LDS
constexpr int NSLOT = 64;      // spread of per-replica atomic counters
constexpr int NCOMP = 12;      // x, y, z, O00..O22 per particle
constexpr int SCAN_TILE = 2048;

// per-replica counter rows, each NSLOT wide
enum Ctr : int { C_TRY_ROTO = 0, C_ACC_ROTO, C_DBONDS, C_TRY_VOL, C_ACC_VOL, C_EBONDS, C_OVL, NCTR };

// device error bits
enum Err : int { ERR_CELL_CAP = 1, ERR_BOX = 2 };

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
	int sel;           // buffer that holds the current state (0/1)
	int ensemble;
	int min_n;         // smallest cells per side of the grid (4 for sweeps)
	int even;          // sweep grids need an even count per side
	int trial_active;  // an NPT trial is in flight
	int trial_dir;
	double side_min;   // smallest allowed cell side
	double T, disp_max, theta_max, rc2, cosmax, P, rfm;
	double V;          // the volume the reference tracks (MC.c:500)
	long long bonds_base;  // bonded pairs = bonds_base + sum of C_DBONDS slots
	double trial_f, trial_V, trial_V0;
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
	int *count, *loc, *start, *slot, *cellid, *idx, *tile_sum;
	int ntiles;
	unsigned long long *ctr;  // [R][NCTR][NSLOT]
	int *err;
	int32_t *out_bonds;
	uint8_t *out_ovl;
	long long *trace;  // [R][trace_cap] bonded pairs at the start of each sweep (optional)
	int trace_cap;
	Base base;
	uint64_t seed;
};

// launchers (pp_kernels.hip)
void launch_fill_i32(int *p, int v, long long n, hipStream_t s);
void launch_fill_u64(unsigned long long *p, unsigned long long v, long long n, hipStream_t s);
void launch_prep(const DevState &d, uint64_t sweep, long long trace_pos, hipStream_t s);
void launch_flip(const DevState &d, hipStream_t s);
// trial: 0 = rebuild every replica on its grid g; 1 = NPT trial rebuild
void launch_rebuild(const DevState &d, int trial, hipStream_t s);
void launch_pass(const DevState &d, uint64_t sweep, int pass, int T, int block, hipStream_t s);
size_t pass_lds_bytes(int T, int block);  // dynamic LDS of one pass workgroup
constexpr size_t MAX_LDS_BYTES = 65536;
void launch_energy(const DevState &d, int trial, int per_particle, int T, hipStream_t s);
void launch_vol_propose(const DevState &d, uint64_t sweep, hipStream_t s);
void launch_vol_decide(const DevState &d, uint64_t sweep, hipStream_t s);
void launch_gram_schmidt(const DevState &d, hipStream_t s);

}  // namespace ppd
