// pp_kernels.hip -- gfx950 kernels of the Kern-Frenkel rototranslation hot path.
//
// Data layout in HBM (per context, all replicas concatenated):
//   f[2][12][stride]  double  x, y, z, O00..O22 of slot j (SoA, two buffers)
//   pid[2][stride]    int     particle index of slot j
//   start[C+1]        int     first slot of each cell, over the concatenated
//                             cell regions of all replicas (one region per
//                             replica, sized for its largest grid)
// Slots are grouped by cell at every rebuild (order inside a cell is
// whatever the atomics gave); the colour pass ranks a cell's members by
// particle index before it draws from them, so the trajectory depends only
// on the state and the seed, never on thread timing.
//
// Sweep = rebuild on a randomly shifted grid + 8 colour passes in random
// order.  In a pass every cell of one colour (i,j,k parities) is active; its
// neighbourhood holds no other active cell, so one team of T lanes per active
// cell can move that cell's particles while all others stay put.  A trial that
// leaves the cell is rejected, which keeps cell populations fixed during the
// pass: the team makes exactly as many sequential trials as the cell holds,
// so a sweep is exactly N attempts, as a reference sweep (MC.c:19-29).
#pragma clang fp contract(off)

#include "pp_device.hpp"

namespace ppd {

// ------------------------------------------------------------------------
// fills

__global__ void k_fill_i32(int *p, int v, long long n) {
	long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, long long n) {
	long long i = (long long) blockIdx.x * blockDim.x + threadIdx.x;
	if(i < n) p[i] = v;
}

void launch_fill_i32(int *p, int v, long long n, hipStream_t s) {
	if(n <= 0) return;
	hipLaunchKernelGGL(k_fill_i32, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, p, v, n);
}

void launch_fill_u64(unsigned long long *p, unsigned long long v, long long n, hipStream_t s) {
	if(n <= 0) return;
	hipLaunchKernelGGL(k_fill_u64, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, p, v, n);
}

// Zero the energy-kernel rows (C_EBONDS, C_OVL) of every replica (nothing
// once a run-ahead sequence of box attempts has stopped: *stop >= 0).
__global__ void k_zero_energy_rows(unsigned long long *ctr, int R, const int *stop) {
	if(stop && *stop >= 0) return;
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= R * 2 * NSLOT) return;
	int r = i / (2 * NSLOT), k = i % (2 * NSLOT);
	int row = k < NSLOT ? C_EBONDS : C_OVL;
	*ctr_at(ctr, r, row, k % NSLOT) = 0ull;
}

// ------------------------------------------------------------------------
// rebuild: count -> scan (flips the buffer selector) -> scatter
//
// mode REB_KEEP   regroup on the current grid and shift
//      REB_SHIFT  new random grid shift for this sweep (uniform in one cell
//                 side per axis), recorded by the first slot of each replica
//      REB_TRIAL  NPT trial: the replicas with a trial in flight are counted
//                 on their trial grid with the trial scaling of one axis
//                 (MC.c:475-478); the others on their grid, unscaled, so that
//                 every replica's cells still start at its own first slot

__device__ __forceinline__ void sweep_shift(const DevState &d, int r, const Grid &G, uint64_t sweep, double sh[3]) {
	u32x4 v = rng(d.seed, r, sweep, 0, TAG_SHIFT, 0, 0);
	sh[0] = u01(v.x) * (G.box[0] / (double) G.n[0]);
	sh[1] = u01(v.y) * (G.box[1] / (double) G.n[1]);
	sh[2] = u01(v.z) * (G.box[2] / (double) G.n[2]);
}

// Replica whose cell region holds global cell c.
__device__ __forceinline__ int replica_of_cell(const DevState &d, int c) {
	int lo = 0, hi = d.R;
	while(hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if(d.reps[mid].cell_off <= c) lo = mid;
		else hi = mid;
	}
	return lo;
}

// A trial that regroups the replica's particles into the spare buffer (a
// box trial whose grid the current lists cannot serve).
__device__ __forceinline__ bool trial_rebuilds(const Rep &rp) { return rp.trial_active && !rp.trial_fast; }

// ONE: a context of one replica, whose table then loads to scalar registers.
template <bool ONE>
__global__ void k_count(DevState d, int mode, uint64_t sweep, long long trace_pos) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	// reset the scan's look-back state for the scan that follows
	for(int t = j; t < d.ntiles; t += gridDim.x * blockDim.x) d.scan_status[t] = 0ull;
	for(int t = j; t < d.R * 8; t += gridDim.x * blockDim.x) d.ccount[t] = 0;
	if(j == 0) {
		d.scan_ticket[0] = 0;
	}
	if(mode == REB_TRIAL)
		// replicas without a trial rebuild keep their lists as they are:
		// their cells' counts are the current populations, so the scan
		// reproduces their starts
		for(int c = j; c < d.C_total; c += gridDim.x * blockDim.x)
			if(!trial_rebuilds(d.reps[replica_of_cell(d, c)])) d.count[c] = d.start[c + 1] - d.start[c];
	if(j >= d.Ntot) return;
	const int r = ONE ? 0 : d.prep[j];
	Rep &rp = d.reps[r];
	const bool tr = (mode == REB_TRIAL) && trial_rebuilds(rp);
	if(mode == REB_TRIAL && !tr) return;
	const Grid &G = tr ? rp.tg : rp.g;
	double sh[3] = {G.shift[0], G.shift[1], G.shift[2]};
	if(mode == REB_SHIFT) {
		sweep_shift(d, r, G, sweep, sh);
		if(j == rp.p_off) {
			// energy trace: bonded pairs as they stand after the previous sweep
			if(d.trace && trace_pos >= 0 && trace_pos < d.trace_cap) {
				unsigned long long db = 0;
				for(int k = 0; k < NSLOT; k++) db += *ctr_at(d.ctr, r, C_DBONDS, k);
				d.trace[(long long) r * d.trace_cap + trace_pos] = rp.bonds_base + (long long) db;
			}
			rp.g.shift[0] = sh[0];
			rp.g.shift[1] = sh[1];
			rp.g.shift[2] = sh[2];
			// random colour order of this sweep's 8 passes
			u32x4 v = rng(d.seed, r, sweep, 0, TAG_COLOUR, 0, 0);
			const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
			rp.col_perm = colour_perm_from(w4);
		}
	}
	if(!tr && j == rp.p_off) {
		// the lists the scan and scatter build are by membership in this grid
		rp.lg = rp.g;
		for(int k = 0; k < 3; k++) rp.lg.shift[k] = sh[k];
		rp.lists_ok = 1;
		rp.rec_ok = mode == REB_SHIFT;  // only that rebuild writes the cell records
	}
	// a free or removed slot of a grand-canonical replica goes to the
	// replica's parking cell
	int c = rp.cell_off + rp.cell_cap;
	if(j - rp.p_off < rp.np_hw && pick_buf(d.pid, rp.sel)[j] >= 0) {
		const double *F = pick_buf(d.f, rp.sel);
		const int S = d.stride;
		double p[3] = {F[j], F[S + j], F[2 * S + j]};
		if(tr)
			for(int k = 0; k < 3; k++) p[k] *= rp.trial_s[k];
		int ix = cell_coord(p[0], sh[0], G.inv_box[0], G.n[0]);
		int iy = cell_coord(p[1], sh[1], G.inv_box[1], G.n[1]);
		int iz = cell_coord(p[2], sh[2], G.inv_box[2], G.n[2]);
		c = rp.cell_off + (ix * G.n[1] + iy) * G.n[2] + iz;
	}
	d.cellid[j] = c;
	d.slot[j] = atomicAdd(&d.count[c], 1);
}

// (replica, colour) keys of the non-empty-cell lists counted per scan workgroup
constexpr int SCAN_KEYS = 512;

// A lane's PER consecutive cells of a tile: whole vectors where all PER
// are in the grid (base is a multiple of PER, the arrays 16-byte aligned).
template <int PER>
__device__ __forceinline__ void load_cells(const int *p, int base, int n, int v[PER]) {
	static_assert(PER == 2 || PER % 4 == 0, "2 or a multiple of 4 cells per lane");
	if(base + PER <= n) {
		if constexpr(PER == 2) {
			const int2 q = *(const int2 *) (p + base);
			v[0] = q.x;
			v[1] = q.y;
		}
		else
#pragma unroll
			for(int e = 0; e < PER; e += 4) {
				const int4 q = *(const int4 *) (p + base + e);
				v[e] = q.x;
				v[e + 1] = q.y;
				v[e + 2] = q.z;
				v[e + 3] = q.w;
			}
	}
	else
#pragma unroll
		for(int e = 0; e < PER; e++) v[e] = base + e < n ? p[base + e] : 0;
}

template <int PER>
__device__ __forceinline__ void store_cells(int *p, int base, int n, const int v[PER]) {
	if(base + PER <= n) {
		if constexpr(PER == 2) *(int2 *) (p + base) = make_int2(v[0], v[1]);
		else
#pragma unroll
			for(int e = 0; e < PER; e += 4) *(int4 *) (p + base + e) = make_int4(v[e], v[e + 1], v[e + 2], v[e + 3]);
	}
	else
#pragma unroll
		for(int e = 0; e < PER; e++)
			if(base + e < n) p[base + e] = v[e];
}

// Exclusive scan of the cell counts in one pass (decoupled look-back):
// workgroups take tiles in dispatch order from a ticket, publish their tile
// sum, then read their predecessors' published values, 64 at a time, back to
// the nearest one that carries an inclusive prefix.  A workgroup only ever waits for tiles with a
// smaller ticket, which are already running.  start[c] = first slot of cell
// c; count is cleared for the next rebuild; a flipping scan also switches
// every replica to the buffer the scatter fills.
// Sum of the cell counts of each scan tile.
template <int PER>
__global__ __launch_bounds__(256) void k_tile_sums(DevState d) {
	__shared__ int part[4];
	const int t = threadIdx.x, base = blockIdx.x * (256 * PER) + t * PER;
	int v[PER];
	load_cells<PER>(d.count, base, d.C_total, v);
	int tot = 0;
#pragma unroll
	for(int e = 0; e < PER; e++) tot += v[e];
	tot = team_sum<64>(tot);
	if((t & 63) == 0) part[t >> 6] = tot;
	__syncthreads();
	if(t == 0) d.tile_sum[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive prefix of the tile sums, in place (one workgroup).
__global__ __launch_bounds__(1024) void k_tile_scan(DevState d) {
	__shared__ int part[1024];
	const int t = threadIdx.x, n = d.ntiles;
	const int per = (n + 1023) / 1024, lo = t * per, hi = min(n, lo + per);
	int s = 0;
	for(int i = lo; i < hi; i++) s += d.tile_sum[i];
	part[t] = s;
	__syncthreads();
	for(int o = 1; o < 1024; o <<= 1) {
		const int v = t >= o ? part[t - o] : 0;
		__syncthreads();
		part[t] += v;
		__syncthreads();
	}
	int run = part[t] - s;
	for(int i = lo; i < hi; i++) {
		const int v = d.tile_sum[i];
		d.tile_sum[i] = run;
		run += v;
	}
}

// pre = 1: the tiles' exclusive prefixes are in d.tile_sum (k_tile_sums,
// k_tile_scan); pre = 0: a decoupled look-back finds them.
// PER cells per lane: tiles of 256 * PER cells.
// ONE: a context of one replica (its table in scalar registers).
template <int PER, bool ONE>
__global__ __launch_bounds__(256) void k_scan(DevState d, int flip, int lists, int pre) {
	__shared__ int wsum[4];
	__shared__ int sh_tile, sh_excl, sh_r0;
	__shared__ int lcnt[SCAN_KEYS], lbase[SCAN_KEYS];
	const int t = threadIdx.x;
	if(t == 0) sh_tile = atomicAdd(d.scan_ticket, 1);
	for(int k = t; k < SCAN_KEYS; k += 256) lcnt[k] = 0;
	__syncthreads();
	const int tile = sh_tile;
	const int base = tile * (256 * PER) + t * PER;
	if(t == 0 && lists && !ONE) {
		// replica of the tile's first cell
		int lo = 0, hi = d.R;
		while(hi - lo > 1) {
			int mid = (lo + hi) >> 1;
			if(d.reps[mid].cell_off <= tile * (256 * PER)) lo = mid;
			else hi = mid;
		}
		sh_r0 = lo;
	}
	int v[PER];
	load_cells<PER>(d.count, base, d.C_total, v);
	int tot = 0;
#pragma unroll
	for(int e = 0; e < PER; e++) tot += v[e];
	int lane = t & 63, wid = t >> 6;
	const int inc = wave_incl_scan(tot);
	if(lane == 63) wsum[wid] = inc;
	__syncthreads();
	if(pre) {
		if(t == 0) sh_excl = d.tile_sum[tile];
	}
	else if(wid == 0) {
		// wavefront 0 publishes the tile's sum and looks back over up to 64
		// predecessors at once: the nearest one with an inclusive prefix ends
		// the walk; one that has published nothing yet is polled again.  A
		// status word carries its flag and its value together, so relaxed
		// agent-scope atomics suffice (a release would write the XCD's L2
		// back on every publication).
		const int agg = wsum[0] + wsum[1] + wsum[2] + wsum[3];
		int excl = 0;
		if(lane == 0)
			__hip_atomic_store(&d.scan_status[tile], ((tile == 0 ? 2ull : 1ull) << 32) | (unsigned) agg,
			                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		int p = tile - 1, spins = 0;
		while(p >= 0) {
			const int q = p - lane;
			unsigned long long st = 2ull << 32;  // before tile 0: an inclusive prefix of 0
			if(q >= 0) st = __hip_atomic_load(&d.scan_status[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			const unsigned flag = (unsigned) (st >> 32);
			const unsigned long long m2 = __ballot(flag == 2u);
			const int l2 = m2 ? __ffsll((long long) m2) - 1 : 64;
			const unsigned long long upto = l2 >= 63 ? ~0ull : ((2ull << l2) - 1ull);
			if(__ballot(flag == 0u) & upto) {
				if(++spins > SCAN_SPIN_MAX) {
					if(lane == 0) atomicOr(d.err, (int) ERR_SCAN);
					break;
				}
				__builtin_amdgcn_s_sleep(1);
				continue;
			}
			int val = lane <= l2 ? (int) (st & 0xffffffffull) : 0;
			excl += team_sum<64>(val);
			if(l2 < 64) break;
			p -= 64;
		}
		if(lane == 0) {
			if(tile > 0)
				__hip_atomic_store(&d.scan_status[tile], (2ull << 32) | (unsigned) (excl + agg), __ATOMIC_RELAXED,
				                   __HIP_MEMORY_SCOPE_AGENT);
			sh_excl = excl;
		}
	}
	__syncthreads();
	int woff = 0;
	for(int w = 0; w < wid; w++) woff += wsum[w];
	// the cells' first slots; their counts cleared for the next rebuild
	{
		int first[PER], zero[PER];
		int run = sh_excl + woff + inc - tot;
#pragma unroll
		for(int e = 0; e < PER; e++) {
			first[e] = run;
			zero[e] = 0;
			run += v[e];
		}
		store_cells<PER>(d.start, base, d.C_total, first);
		store_cells<PER>(d.count, base, d.C_total, zero);
	}
	int lkey[PER], lpos[PER];
#pragma unroll
	for(int e = 0; e < PER; e++) {
		int c = base + e;
		lkey[e] = -1;
		lpos[e] = 0;
		if(c < d.C_total) {
			if(lists && v[e] > 0) {
				// the next passes visit only non-empty cells: list them by
				// colour, first counted per workgroup in LDS
				int lo = ONE ? 0 : sh_r0, hi = d.R;
				if constexpr(!ONE)
					while(hi - lo > 1) {
						int mid = (lo + hi) >> 1;
						if(d.reps[mid].cell_off <= c) lo = mid;
						else hi = mid;
					}
				const Rep &rp = d.reps[ONE ? 0 : lo];
				const int cloc = c - rp.cell_off;
				const int n1 = rp.g.n[1], n2 = rp.g.n[2];
				const int ix = cloc / (n1 * n2), iy = (cloc / n2) % n1, iz = cloc % n2;
				const int col = ((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1);
				const int key = ONE ? col : (lo - sh_r0) * 8 + col;
				if(cloc >= rp.g.n[0] * n1 * n2) {
					// the parking cell: free slots, no particles to move
				}
				else if(key < SCAN_KEYS) {
					lkey[e] = key;
					lpos[e] = atomicAdd(&lcnt[key], 1);
				}
				else {
					// a tile spanning more than SCAN_KEYS / 8 replicas: direct
					const int k = atomicAdd(&d.ccount[lo * 8 + col], 1);
					d.clist[rp.cell_off + col * rp.act_cap + k] = cloc;
				}
			}
		}
	}
	if(lists) {
		__syncthreads();
		for(int k = t; k < (ONE ? 8 : SCAN_KEYS); k += 256) {
			const int n = lcnt[k];
			if(n > 0) lbase[k] = atomicAdd(&d.ccount[((ONE ? 0 : sh_r0) + k / 8) * 8 + k % 8], n);
		}
		__syncthreads();
#pragma unroll
		for(int e = 0; e < PER; e++)
			if(lkey[e] >= 0) {
				const int r = ONE ? 0 : sh_r0 + lkey[e] / 8;
				const Rep &rp = d.reps[r];
				const int cloc = base + e - rp.cell_off;
				d.clist[rp.cell_off + (lkey[e] & 7) * rp.act_cap + lbase[lkey[e]] + lpos[e]] = cloc;
			}
	}
	if(tile == 0) {
		if(t == 0) d.start[d.C_total] = d.Ntot;
		if(flip)
			for(int r = t; r < d.R; r += 256) d.reps[r].sel ^= 1;
	}
}

// The layout of the passes' lists, after the scan has counted them:
// pfx[k][r] = first entry of replica r in the list of pass k (one thread per
// pass; the kernel boundary makes every workgroup's counts visible).  Also
// resets the ticket queue of the dependency-driven sweep.
__device__ __forceinline__ void pass_prefix(const DevState &d, int t) {
	if(t == 0 && d.ddq) *d.ddq = 0;  // the dependency-driven sweep's tickets
	if(t < 8) {
		if(d.ovf_n) d.ovf_n[t] = 0;  // the small-cell pass's deferred cells
		int acc = 0;
		d.pfx[t * (d.R + 1)] = 0;
		for(int r = 0; r < d.R; r++) {
			const int col = (int) ((d.reps[r].col_perm >> (4 * t)) & 0xFu);
			acc += d.ccount[r * 8 + col];
			d.pfx[t * (d.R + 1) + r + 1] = acc;
		}
	}
}

__global__ void k_pass_prefix(DevState d) { pass_prefix(d, threadIdx.x); }

// Record of entry i of the cell lists (replica r's list of colour col at
// clist[cell_off + col * act_cap + k]), stored at the entry's position in the
// list of the pass that works on that colour: the 26 surrounding cells' first
// slots [0, 26), the exclusive prefix of their populations [32, 59), the
// cell's first slot, population, local index, replica and buffer selector
// [59, 64).  A pass finds all its data from this one load.
template <bool ONE>
__device__ __forceinline__ void build_record(const DevState &d, int i) {
	int lo = 0, hi = d.R;
	if constexpr(!ONE)
		while(hi - lo > 1) {
			int mid = (lo + hi) >> 1;
			if(d.reps[mid].cell_off <= i) lo = mid;
			else hi = mid;
		}
	const Rep &rp = d.reps[lo];
	const int local = i - rp.cell_off;
	const int col = local / rp.act_cap, kk = local - col * rp.act_cap;
	if(col >= 8 || kk >= d.ccount[lo * 8 + col]) return;
	const int cloc = d.clist[i];
	int k = 0;
	while(k < 7 && (int) ((rp.col_perm >> (4 * k)) & 0xFu) != col) k++;
	const int gi = (ONE ? 0 : d.pfx[k * (d.R + 1) + lo]) + kk;
	const int n0 = rp.g.n[0], n1 = rp.g.n[1], n2 = rp.g.n[2];
	const int ix = cloc / (n1 * n2), iy = (cloc / n2) % n1, iz = cloc % n2;
	const int *st = d.start + rp.cell_off;
	int rec[REC_W];
	int acc = 0;
#pragma unroll
	for(int nb = 0; nb < 26; nb++) {
		const int m = nb < 13 ? nb : nb + 1;
		const int cx = wrapi(ix + m / 9 - 1, n0), cy = wrapi(iy + (m / 3) % 3 - 1, n1), cz = wrapi(iz + m % 3 - 1, n2);
		const int cc = (cx * n1 + cy) * n2 + cz;
		const int a = st[cc], b = st[cc + 1];
		rec[nb] = a;
		rec[32 + nb] = acc;
		acc += b - a;
	}
#pragma unroll
	for(int e = 26; e < 32; e++) rec[e] = 0;
	const int s0 = st[cloc];
	rec[58] = acc;
	rec[59] = s0;
	rec[60] = st[cloc + 1] - s0;
	rec[61] = cloc;
	rec[62] = lo;
	rec[63] = rp.sel;
	int4 *out = (int4 *) (d.rec + ((size_t) k * d.total_act + gi) * REC_W);
#pragma unroll
	for(int e = 0; e < REC_W / 4; e++) out[e] = make_int4(rec[4 * e], rec[4 * e + 1], rec[4 * e + 2], rec[4 * e + 3]);
}

// Move every particle to its cell's slot range.  After a flipping scan the
// source is the previous buffer (sel ^ 1); a trial writes into sel ^ 1.
// With a new grid shift the workers also write the cell records of the
// sweep's passes.
// ONE: one replica, whose lists start at entry 0 of every pass; the first
// workgroup also writes the pass prefixes (k_pass_prefix's work), which the
// records then do not need.
template <bool ONE>
__global__ void k_scatter(DevState d, int mode, uint64_t sweep) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if constexpr(ONE)
		if(mode == REB_SHIFT && blockIdx.x == 0) pass_prefix(d, threadIdx.x);
	// the particle's loads first, so that they overlap the record's chain of
	// loads; its stores after the record's
	bool mv = false, hw = false;
	int r = 0, dst = 0, pid = 0;
	double v[NCOMP];
	double *G = nullptr;
	int *PG = nullptr;
	const int S = d.stride;
	if(j < d.Ntot) {
		r = ONE ? 0 : d.prep[j];
		const Rep &rp = d.reps[r];
		const bool tr = (mode == REB_TRIAL) && trial_rebuilds(rp);
		if(mode != REB_TRIAL || tr) {
			mv = true;
			// the buffer this rebuild switches to holds the live particles first
			hw = mode != REB_TRIAL && j == rp.p_off;
			const int sb = tr ? rp.sel : (rp.sel ^ 1);
			const double *F = pick_buf(d.f, sb);
			G = pick_buf(d.f, sb ^ 1);
			PG = pick_buf(d.pid, sb ^ 1);
			dst = d.start[d.cellid[j]] + d.slot[j];
#pragma unroll
			for(int c = 0; c < NCOMP; c++) {
				v[c] = F[c * S + j];
				if(tr && c < 3) v[c] *= rp.trial_s[c];
			}
			pid = pick_buf(d.pid, sb)[j];
		}
	}
	if(mode == REB_SHIFT && j < d.C_total) build_record<ONE>(d, j);
	if(!mv) return;
#pragma unroll
	for(int c = 0; c < NCOMP; c++) G[c * S + dst] = v[c];
	PG[dst] = pid;
	if(hw) d.reps[r].np_hw = d.reps[r].np;
}

void launch_rebuild(const DevState &d, int mode, uint64_t sweep, long long trace_pos, hipStream_t s) {
	const int B = 256;
	unsigned gp = (unsigned) ((d.Ntot + B - 1) / B);
	if(gp == 0) gp = 1;
	const int flip = mode != REB_TRIAL;
	if(d.R == 1) hipLaunchKernelGGL(k_count<true>, dim3(gp), dim3(B), 0, s, d, mode, sweep, trace_pos);
	else hipLaunchKernelGGL(k_count<false>, dim3(gp), dim3(B), 0, s, d, mode, sweep, trace_pos);
	const int lists = mode == REB_SHIFT ? 1 : 0;
	if(d.scan_tile == SCAN_TILE_SMALL) {
		// at most SCAN_SMALL_MAX_TILES tiles: one pass with a look-back
		if(d.R == 1) hipLaunchKernelGGL((k_scan<SCAN_TILE_SMALL / 256, true>), dim3(d.ntiles), dim3(256), 0, s, d, flip, lists, 0);
		else hipLaunchKernelGGL((k_scan<SCAN_TILE_SMALL / 256, false>), dim3(d.ntiles), dim3(256), 0, s, d, flip, lists, 0);
	}
	else {
		// large tiles are taken only beyond SCAN_SMALL_MAX_TILES small ones,
		// so they always come in more tiles than a look-back serves well:
		// tile sums, one scan of the sums, then the tiles
		static_assert(SCAN_SMALL_MAX_TILES * SCAN_TILE_SMALL / SCAN_TILE_LARGE > SCAN_LOOKBACK_MAX_TILES,
		              "grids of large tiles always exceed the look-back's tile limit");
		hipLaunchKernelGGL(k_tile_sums<SCAN_TILE_LARGE / 256>, dim3(d.ntiles), dim3(256), 0, s, d);
		hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, d);
		hipLaunchKernelGGL((k_scan<SCAN_TILE_LARGE / 256, false>), dim3(d.ntiles), dim3(256), 0, s, d, flip, lists, 1);
	}
	if(mode == REB_SHIFT && d.R != 1) hipLaunchKernelGGL(k_pass_prefix, dim3(1), dim3(64), 0, s, d);
	unsigned gs = gp;
	if(mode == REB_SHIFT) gs = std::max(gp, (unsigned) ((d.C_total + B - 1) / B));
	if(d.R == 1) hipLaunchKernelGGL(k_scatter<true>, dim3(gs), dim3(B), 0, s, d, mode, sweep);
	else hipLaunchKernelGGL(k_scatter<false>, dim3(gs), dim3(B), 0, s, d, mode, sweep);
}

// ------------------------------------------------------------------------
// the colour pass

struct PassArgs {
	DevState d;
	uint64_t sweep;
	int pass;
	unsigned mark;  // dependency-driven sweeps: the value a finished cell publishes in done[]
};

// Per-team LDS (doubles) for cells of up to CAP members: the stage row of a
// trial's O' and patches, the cell's trials, its members ranked by particle
// index (x, y, z, O and the 4 patch vectors; the kept form holds the
// member's fractional cell coordinates in the first three patch slots),
// their particle indices, their kept static bond counts and the cell's
// record (surrounding slot ranges).
constexpr int OWN_W = NCOMP + 12;   // x, y, z, O[9], P[12]
constexpr int LDS_NB = REC_W / 2;   // the cell record, REC_W ints
constexpr int LDS_STAGE = 22;       // a trial's O' (9) and patches (12), formed one value per lane
template <int CAP>
struct TeamLds {
	static constexpr int OWN = CAP * OWN_W;
	static constexpr int PID = (CAP + 1) / 2;  // CAP ints
	static constexpr int SB = (CAP + 1) / 2;   // CAP ints: each member's bonds with the static neighbours
	// a trial per member, or a fixed number per cell
	static constexpr int TRIAL = (CAP > MAX_TRIALS_PER_CELL ? CAP : MAX_TRIALS_PER_CELL) * TRIAL_W;
	static constexpr int FIXED = OWN + PID + SB + TRIAL + LDS_NB + LDS_STAGE;
	static_assert((LDS_STAGE + TRIAL + OWN + PID + SB) % 2 == 0 && FIXED % 2 == 0 && (LDS_STAGE + TRIAL) % 2 == 0 &&
	                  TRIAL_W % 2 == 0 && OWN_W % 2 == 0,
	              "16-byte aligned trials, members, record and team blocks");
};
// Cells of the small-cell pass (dilute systems, 16-lane teams with one
// register slot per lane): about half the LDS of a general team, so twice
// the cells stay resident; fuller cells are deferred to the general pass.
constexpr int SMALL_CAP = 8;
// + the orientations of the register neighbours, component-major (two
// slots per lane for 16-lane teams, one per lane for the small-cell pass)
template <int T, int CAP = CELL_CAP>
constexpr int team_lds() {
	return TeamLds<CAP>::FIXED + 9 * (CAP < CELL_CAP ? T : (T < 32 ? 32 : T));
}
// 16 teams of 32 lanes (four 128-lane workgroups, the default shape) share a
// CU's 160 KiB; 32 small-cell teams of 16 lanes do
static_assert(2 * 8 * team_lds<32>() * sizeof(double) <= 160 * 1024, "two pass workgroups per CU");
static_assert(4 * 8 * team_lds<16, SMALL_CAP>() * sizeof(double) <= 160 * 1024, "four small-cell workgroups per CU");

// The same decisions through the filtered tests (no square root, no
// divisions, no branch per patch).  A pair inside a filter band sets `band`;
// the caller then redoes that lane's neighbours with bonds_old_new_exact.
template <class LoadO>
__device__ __forceinline__ void bonds_old_new(const Trial &t, const double *Po, const double *Pn, double qx, double qy,
                                              double qz, LoadO load_o, int &bo, int &bn, bool &band) {
	const double ox = kf_min_image(qx - t.px, t.Lx, t.iLx);
	const double oy = kf_min_image(qy - t.py, t.Ly, t.iLy);
	const double oz = kf_min_image(qz - t.pz, t.Lz, t.iLz);
	const double o2 = ox * ox + oy * oy + oz * oz;
	const double wx = kf_min_image(qx - t.nx, t.Lx, t.iLx);
	const double wy = kf_min_image(qy - t.ny, t.Ly, t.iLy);
	const double wz = kf_min_image(qz - t.nz, t.Lz, t.iLz);
	const double w2 = wx * wx + wy * wy + wz * wz;
	const KfBand bo_ = kf_band(t.c2, o2), bn_ = kf_band(t.c2, w2);
	const unsigned fo = o2 < t.rc2 ? kf_fast_p(ox, oy, oz, Po, bo_) : 0u;
	const unsigned fn = w2 < t.rc2 ? kf_fast_p(wx, wy, wz, Pn, bn_) : 0u;
	unsigned unsure = ((fo | fn) >> 1) & 1u;
	if((fo | fn) & 1u) {
		double Oq[9];
		load_o(Oq);
		unsigned qo, qn;
		kf_fast_q2(ox, oy, oz, kf_band_q(bo_), wx, wy, wz, kf_band_q(bn_), Oq, qo, qn);
		qo = (fo & 1u) ? qo : 0u;
		qn = (fn & 1u) ? qn : 0u;
		unsure |= ((qo | qn) >> 1) & 1u;
		bo += (int) (qo & 1u);
		bn += (int) (qn & 1u);
	}
	band |= unsure != 0u;
}

// One cell of a colour pass: entry gi of the pass's list of non-empty cells
// (all replicas), worked on by the T lanes of one team with its LDS block
// `own`.  The cell's members
// make sequential trials.  Each lane keeps up to CPL of the cell's static
// neighbours (the 26 surrounding cells, fixed during the pass) in registers;
// a denser neighbourhood is handled T at a time from HBM.
// One of the 21 values of a trial's O' = O R (v < 9: O'[v]) and its patches
// (v = 9 + 3a + kk: P'[a][kk]): the row of O it needs, the column, and for a
// patch component the signs of patch a as sign-bit masks of the high word.
struct RotShare {
	int row, col;
	bool patch;
	unsigned long long s0, s1, s2;
};

__device__ __forceinline__ RotShare rot_share(int v) {
	RotShare r;
	const int idx = v < 9 ? 0 : v - 9;
	r.patch = v >= 9;
	r.row = v < 9 ? v / 3 : (v < 21 ? idx % 3 : 0);
	r.col = v % 3;
	const unsigned neg = (v >= 9 && v < 21) ? (KF_NEG >> (3 * (idx / 3))) & 7u : 0u;
	r.s0 = (unsigned long long) (neg & 1u) << 63;
	r.s1 = (unsigned long long) ((neg >> 1) & 1u) << 63;
	r.s2 = (unsigned long long) ((neg >> 2) & 1u) << 63;
	return r;
}

// The value: row `row` of O' from that row of O (o0, o1, o2) and R (the
// trial's tv[4..12]) in the reference's order (utils.c:172-184), then either
// its column or, with b_a[j] = +-c and (+-x) c = +-(x c) exactly, the signed
// sum of the row times c (utils.h:17-21).
__device__ __forceinline__ double rot_value(const RotShare &r, double o0, double o1, double o2, const double *tv) {
	const double r0 = o0 * tv[4] + o1 * tv[4 + 3] + o2 * tv[4 + 6];
	const double r1 = o0 * tv[5] + o1 * tv[5 + 3] + o2 * tv[5 + 6];
	const double r2 = o0 * tv[6] + o1 * tv[6 + 3] + o2 * tv[6 + 6];
	if(!r.patch) return r.col == 0 ? r0 : (r.col == 1 ? r1 : r2);
	const double m0 = r0 * KF_C, m1 = r1 * KF_C, m2 = r2 * KF_C;
	const double a0 = __longlong_as_double(__double_as_longlong(m0) ^ (long long) r.s0);
	const double a1 = __longlong_as_double(__double_as_longlong(m1) ^ (long long) r.s1);
	const double a2 = __longlong_as_double(__double_as_longlong(m2) ^ (long long) r.s2);
	return a0 + a1 + a2;
}

// Trials per cell from which the kept form pays: with fewer, the up-front
// counts cost about what the trials save.
constexpr int KEEP_MIN_TRIALS = 4;

// Margin of the reach filter over rounding in positions and distances
// (below 1e-12 for coordinates of a few hundred).
constexpr double REACH_EPS = 1e-9;

// A read of another cell's particle data during a pass.  In the
// dependency-driven sweep (DD: the passes of a sweep in one launch, each cell
// starting once its neighbours of earlier passes are done) it is an
// agent-scope load, served where the writes of every XCD meet, so that no
// XCD's L2 can return a copy older than the neighbour's finished pass.
template <bool DD>
__device__ __forceinline__ double nbld(const double *p) {
	if constexpr(DD) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	else return *p;
}

// Wait (DD) until every non-empty neighbour cell that an earlier pass of this
// sweep works on has published `mark` in done[]: lane nb < 26 of the team
// watches neighbour nb of the cell record (direction m = nb < 13 ? nb : nb +
// 1, offsets m / 9 - 1, (m / 3) % 3 - 1, m % 3 - 1).  Neighbours of later
// passes wait for this cell in turn, so none of them writes before this
// cell's reads.  The wait is bounded: past DD_SPIN_MAX polls the team reports
// ERR_DEPWAIT and goes on, so the launch always drains.
template <int T>
__device__ __forceinline__ void wait_neighbours(const DevState &d, const Rep &rp, const int *nb_tab, int pass, int ix,
                                                int iy, int iz, int n0, int n1, int n2, unsigned mark, int lane) {
	bool need = false;
	const unsigned *w = nullptr;
	if(lane < 26) {
		const int m = lane < 13 ? lane : lane + 1;
		const int cx = wrapi(ix + m / 9 - 1, n0), cy = wrapi(iy + (m / 3) % 3 - 1, n1), cz = wrapi(iz + m % 3 - 1, n2);
		const int ncol = ((cx & 1) << 2) | ((cy & 1) << 1) | (cz & 1);
		int np = 0;
		while(np < 7 && (int) ((rp.col_perm >> (4 * np)) & 0xFu) != ncol) np++;
		need = np < pass && nb_tab[32 + lane + 1] > nb_tab[32 + lane];
		w = d.done + rp.cell_off + (cx * n1 + cy) * n2 + cz;
	}
	int spins = 0;
	while(team_any<T>(need)) {
		if(need) need = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != mark;
		if(++spins > DD_SPIN_MAX) {
			if(lane == 0) atomicOr(d.err, (int) ERR_DEPWAIT);
			break;
		}
		if(team_any<T>(need)) __builtin_amdgcn_s_sleep(1);
	}
	// no load of the neighbours' data moves above the wait
	__atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// KEEP (fixed trials per cell, many per member): every member's bonds with
// the static neighbours are counted up front and kept as it moves, so a trial
// tests only the new position against them.  Without KEEP (about one trial
// per member) every trial tests both positions.
//
// DD: the pass is one of several in flight (k_pass_dd).  The cell's own
// loads and its trials come first; the team then waits for its neighbours of
// earlier passes, reads their data with agent-scope loads, and finally
// writes its members back with agent-scope stores, waits for them to
// complete, and publishes `mark` for its own cell.
//
// CAP < CELL_CAP (the small-cell pass): a cell of more than CAP members is
// not worked on here but listed for the general pass that follows.
template <int T, int CPL, bool KEEP, bool ONE, bool DD, bool REACH = false, int CAP = CELL_CAP>
__device__ __forceinline__ void pass_cell(const DevState &d, uint64_t sweep, int pass, int gi, int lane,
                                          double *blk, unsigned mark) {
	using L = TeamLds<CAP>;
	// the team's block, the per-trial data first (short LDS offsets)
	double *stage = blk;                    // LDS_STAGE doubles
	double *trials = blk + LDS_STAGE;       // L::TRIAL
	double *own = trials + L::TRIAL;        // L::OWN
	int *own_pid = (int *) (own + L::OWN);
	int *sbond = (int *) (own + L::OWN + L::PID);  // member m's bonds with the static neighbours
	int *nb_tab = (int *) (own + L::OWN + L::PID + L::SB);  // the cell's record (build_record)
	// the orientations of this lane's register neighbours: in registers where
	// they allow (scalar replica constants, and one slot per lane or a reach
	// filter that leaves the second slot almost always empty), else in LDS,
	// component c of neighbour u at (9 u + c) * T + lane
	constexpr bool REG_ORI = (ONE || T == 64) && (CPL == 1 || REACH);
	double *ori = blk + L::FIXED;
	const int S = d.stride;
	{
		// 1. the cell's record: surrounding slot ranges, own range, replica
		const int4 *rec = (const int4 *) (d.rec + ((size_t) pass * d.total_act + gi) * REC_W);
		for(int e = lane; e < REC_W / 4; e += T) ((int4 *) nb_tab)[e] = rec[e];
		team_sync();
		const int M = uni<T>(nb_tab[58]);
		const int s0 = uni<T>(nb_tab[59]);
		const int cnt = uni<T>(nb_tab[60]);
		const int cloc = uni<T>(nb_tab[61]);
		const int r = ONE ? 0 : uni<T>(nb_tab[62]);
		const int sel = uni<T>(nb_tab[63]);
		if constexpr(CAP < CELL_CAP)
			if(cnt > d.small_cap) {
				// (entries < total_act: each cell is listed at most once a pass)
				if(lane == 0) d.ovf[atomicAdd(d.ovf_n + pass, 1)] = gi;
				return;
			}
		// (a team of T < CELL_CAP lanes holds one member per lane)
		if(cnt > CELL_CAP || cnt > T) {
			if(lane == 0) {
				atomicOr(d.err, (int) ERR_CELL_CAP);
				// (DD) the neighbours of later passes do not wait for it
				if constexpr(DD)
					__hip_atomic_store(d.done + d.reps[r].cell_off + cloc, mark, __ATOMIC_RELAXED,
					                   __HIP_MEMORY_SCOPE_AGENT);
			}
			return;
		}
		const Rep &rp = d.reps[r];
		double *F = pick_buf(d.f, sel);
		int *PID = pick_buf(d.pid, sel);

		// 2. every load of the cell at once: the members, the static
		// neighbours' positions and the replica's geometry
		int my_pid = 0;
		double mine[NCOMP];
		if(lane < cnt) {
			my_pid = PID[s0 + lane];
#pragma unroll
			for(int c = 0; c < NCOMP; c++) mine[c] = F[c * S + s0 + lane];
		}
		// trials of the cell in this pass: one per member (a sweep is then N
		// attempts), or a fixed number per non-empty cell
		const int ntrial = d.trials_per_cell > 0 ? d.trials_per_cell : cnt;
		// slot of static neighbour i (0 <= i < M)
		auto nb_slot = [&](int i) {
			int lo2 = 0, hi2 = 26;
			while(hi2 - lo2 > 1) {
				int mid = (lo2 + hi2) >> 1;
				if(nb_tab[32 + mid] <= i) lo2 = mid;
				else hi2 = mid;
			}
			return nb_tab[lo2] + (i - nb_tab[32 + lo2]);
		};
		// Register slot lane + u T (u < CPL) of the team holds one neighbour:
		// slots below cnt the cell's members (member = lane, kept current
		// from the cell's LDS block at every trial), then the first static
		// neighbours; the rest (dense neighbourhoods only) are read again at
		// every trial.  A lane thus tests one neighbour per trial whether the
		// cell holds one member or several.
		double cxr[CPL], cyr[CPL], czr[CPL];
		int cqr[CPL];  // the static neighbour's slot, -1: a member or none
		double orv[CPL][9];
		auto load_static = [&]() {
#pragma unroll
			for(int u = 0; u < CPL; u++) {
				const int i = lane + u * T - cnt;
				cqr[u] = -1;
				cxr[u] = cyr[u] = czr[u] = 0.;
				if(i >= 0 && i < M) {
					const int q = nb_slot(i);
					cqr[u] = q;
					cxr[u] = nbld<DD>(F + q);
					cyr[u] = nbld<DD>(F + S + q);
					czr[u] = nbld<DD>(F + 2 * S + q);
				}
			}
#pragma unroll
			for(int u = 0; u < CPL; u++)
#pragma unroll
				for(int c = 0; c < 9; c++) orv[u][c] = cqr[u] >= 0 ? nbld<DD>(F + (3 + c) * S + cqr[u]) : 0.;
		};
		if constexpr(!DD) load_static();
		Trial tr;
		tr.Lx = unid<T>(rp.g.box[0]);
		tr.Ly = unid<T>(rp.g.box[1]);
		tr.Lz = unid<T>(rp.g.box[2]);
		tr.iLx = unid<T>(rp.g.inv_box[0]);
		tr.iLy = unid<T>(rp.g.inv_box[1]);
		tr.iLz = unid<T>(rp.g.inv_box[2]);
		tr.rc2 = unid<T>(rp.rc2);
		tr.cosmax = unid<T>(rp.cosmax);
		tr.c2 = tr.cosmax >= KF_FILTER_MIN_COS ? tr.cosmax * tr.cosmax : -1.;
		const double temp = unid<T>(rp.T);
		const double sh0 = rp.g.shift[0], sh1 = rp.g.shift[1], sh2 = rp.g.shift[2];
		const int n0 = uni<T>(rp.g.n[0]), n1 = uni<T>(rp.g.n[1]), n2 = uni<T>(rp.g.n[2]);
		const int ix = cloc / (n1 * n2), iy = (cloc / n2) % n1, iz = cloc % n2;

		// 3. the cell's trials, one lane each (Philox: no memory traffic),
		// and the particle indices into LDS.  A trial's slot 0 (the member
		// rank) is replaced by its acceptance bits: bit j is set iff the draw
		// u < exp(-j/T), the replica's factor for a rise of j bonds
		// (MC.c:332), so a trial decides without a table read.  Lane t keeps
		// trial t's rank.
		int my_k = 0;
		double my_dl = 0.;  // (REACH) the length of this lane's trial displacement
		if(lane < ntrial) {
			double tv[TRIAL_W];
			const int col = ((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1);
			make_trial(d, rp, r, sweep, cloc, col, lane, cnt, tv);
			if constexpr(REACH) my_dl = __builtin_sqrt(tv[1] * tv[1] + tv[2] * tv[2] + tv[3] * tv[3]);
			unsigned am = 0;
#pragma unroll
			for(int j = 1; j < ACC_TAB; j++) am |= (unsigned) (tv[13] < rp.acc_tab[j]) << j;
			my_k = (int) tv[0];
			tv[0] = (double) am;
#pragma unroll
			for(int e = 0; e < TRIAL_W; e++) trials[lane * TRIAL_W + e] = tv[e];
		}
		static_assert(MAX_TRIALS_PER_CELL <= 32 && CELL_CAP <= 32, "one lane per trial");
		const int koff = (threadIdx.x & 63) & ~(T - 1);
		if(lane < cnt) own_pid[lane] = my_pid;
		if constexpr(!REG_ORI && !DD && !REACH) {
#pragma unroll
			for(int u = 0; u < CPL; u++)
#pragma unroll
				for(int c = 0; c < 9; c++) ori[(9 * u + c) * T + lane] = orv[u][c];
		}
		team_sync();
		// 4. members into LDS ranked by particle index, with their patch vectors
		if(lane < cnt) {
			int rank = 0;
			for(int m = 0; m < cnt; m++) rank += own_pid[m] < my_pid;
			double *w = own + rank * OWN_W;
#pragma unroll
			for(int c = 0; c < NCOMP; c++) w[c] = mine[c];
			double P[4][3];
			kf_patches(mine + 3, P);
#pragma unroll
			for(int i = 0; i < 4; i++)
#pragma unroll
				for(int c = 0; c < 3; c++) w[NCOMP + 3 * i + c] = P[i][c];
		}
		team_sync();
		if constexpr(REACH) {
			// Keep only the static neighbours a member can reach.  No member
			// travels further during the pass than S, the summed lengths of
			// the cell's trial displacements (triangle inequality on the
			// periodic box), so a neighbour further than 1 + delta + S from
			// every member's start can neither overlap nor bond with any
			// trial position: dropping it changes no decision.  The kept ones
			// are compacted into the register slots after the members (most
			// teams then need one slot per lane) through the orientation
			// block of LDS, 8 values then 5 values per slot.
			const double S = team16_sum_f64(my_dl);
			const double reach = __builtin_sqrt(tr.rc2) + S + REACH_EPS;
			const double reach2 = reach * reach;
			bool keep_u[CPL];
#pragma unroll
			for(int u = 0; u < CPL; u++) {
				double dmin2 = reach2;
				if(cqr[u] >= 0)
					for(int m = 0; m < cnt; m++) {
						const double *wm = own + m * OWN_W;
						const double dx = kf_min_image(cxr[u] - wm[0], tr.Lx, tr.iLx);
						const double dy = kf_min_image(cyr[u] - wm[1], tr.Ly, tr.iLy);
						const double dz = kf_min_image(czr[u] - wm[2], tr.Lz, tr.iLz);
						dmin2 = __builtin_fmin(dmin2, dx * dx + dy * dy + dz * dz);
					}
				keep_u[u] = cqr[u] >= 0 && dmin2 < reach2;
			}
			static_assert(!REACH || (T == 16 && CPL == 2), "reach filter: 16-lane teams, two slots per lane");
			const unsigned long long tmask = ((1ull << T) - 1ull) << koff;
			const unsigned long long k0 = __ballot(keep_u[0]) & tmask, k1 = __ballot(keep_u[CPL - 1]) & tmask;
			const int n0 = __popcll(k0), nr = n0 + __popcll(k1);
			const int dst[2] = {cnt + lanes_below(k0), cnt + n0 + lanes_below(k1)};
			double *stg = ori;  // value v of slot j at stg[32 v + j]
			double stale[CPL][5];
#pragma unroll
			for(int u = 0; u < CPL; u++) {
#pragma unroll
				for(int c = 0; c < 5; c++) stale[u][c] = orv[u][4 + c];
				if(keep_u[u]) {
					const int j = dst[u];
					stg[j] = (double) cqr[u];
					stg[32 + j] = cxr[u];
					stg[64 + j] = cyr[u];
					stg[96 + j] = czr[u];
#pragma unroll
					for(int c = 0; c < 4; c++) stg[32 * (4 + c) + j] = orv[u][c];
				}
			}
			team_sync();
#pragma unroll
			for(int u = 0; u < CPL; u++) {
				const int j = lane + u * T;
				const bool in = j >= cnt && j < cnt + nr;
				cqr[u] = in ? (int) stg[j] : -1;
				cxr[u] = in ? stg[32 + j] : 0.;
				cyr[u] = in ? stg[64 + j] : 0.;
				czr[u] = in ? stg[96 + j] : 0.;
#pragma unroll
				for(int c = 0; c < 4; c++) orv[u][c] = in ? stg[32 * (4 + c) + j] : 0.;
			}
			team_sync();
#pragma unroll
			for(int u = 0; u < CPL; u++)
				if(keep_u[u])
#pragma unroll
					for(int c = 0; c < 5; c++) stg[32 * c + dst[u]] = stale[u][c];
			team_sync();
#pragma unroll
			for(int u = 0; u < CPL; u++) {
				const int j = lane + u * T;
				const bool in = j >= cnt && j < cnt + nr;
#pragma unroll
				for(int c = 0; c < 5; c++) orv[u][4 + c] = in ? stg[32 * c + j] : 0.;
			}
			team_sync();
			if constexpr(!REG_ORI) {
#pragma unroll
				for(int u = 0; u < CPL; u++)
#pragma unroll
					for(int c = 0; c < 9; c++) ori[(9 * u + c) * T + lane] = orv[u][c];
			}
			team_sync();
		}
		if constexpr(DD) {
			// the neighbours' data once their earlier passes are done
			wait_neighbours<T>(d, rp, nb_tab, pass, ix, iy, iz, n0, n1, n2, mark, lane);
			load_static();
			if constexpr(!REG_ORI) {
#pragma unroll
				for(int u = 0; u < CPL; u++)
#pragma unroll
					for(int c = 0; c < 9; c++) ori[(9 * u + c) * T + lane] = orv[u][c];
			}
			team_sync();
		}
		// (orientations in LDS) a member slot's orientation, kept current by
		// the trials that move it
		if constexpr(!REG_ORI)
			if(lane < cnt) {
#pragma unroll
				for(int c = 0; c < 9; c++) ori[c * T + lane] = own[lane * OWN_W + 3 + c];
			}
		// the orientation of the neighbour in register slot u
		auto reg_o = [&](int u, double *Oq) {
#pragma unroll
			for(int c = 0; c < 9; c++) Oq[c] = REG_ORI ? orv[u][c] : ori[(9 * u + c) * T + lane];
		};
		// whether any team of the wavefront has neighbours beyond the
		// register slots (a scalar branch around those loops)
		const bool wave_far = __ballot(M + cnt > T * CPL) != 0ull;
		// the static neighbours beyond the register slots, from HBM
		auto each_far = [&](auto &&f) {
			if(wave_far)
				for(int b = T * CPL - cnt; b < M; b += T)
					if(b + lane < M) {
						const int q = nb_slot(b + lane);
						f(nbld<DD>(F + q), nbld<DD>(F + S + q), nbld<DD>(F + 2 * S + q), [&](double *Oq) {
#pragma unroll
							for(int c = 0; c < 9; c++) Oq[c] = nbld<DD>(F + (3 + c) * S + q);
						});
					}
		};
		team_sync();
		// 5. KEEP: every member's bonds with the static neighbours, and with
		// each other
		unsigned mbits = 0;
		if constexpr(KEEP) {
			// each member lane's own member in its first register slot (the
			// trials refresh it as members move)
			const bool mem_lane = lane < cnt;
			if(mem_lane) {
				const double *wl = own + lane * OWN_W;
				cxr[0] = wl[0];
				cyr[0] = wl[1];
				czr[0] = wl[2];
				if constexpr(REG_ORI)
#pragma unroll
					for(int c = 0; c < 9; c++) orv[0][c] = wl[3 + c];
			}
			// Two members at a time (independent chains), one reduction for
			// both.  Each lane tests its register slots against member mm: a
			// static neighbour adds to the member's kept count; another
			// member sets a bit of the lane's mbits (bit mm of lane j iff
			// members j and mm are bonded: the pair decision is symmetric).
			auto slot_for = [&](int u, int mm) { return u == 0 && mem_lane ? lane != mm : cqr[u] >= 0; };
			auto add = [&](int u, int b, int &bs, int &bm) {
				if(u == 0 && mem_lane) bm = b;
				else bs += b;
			};
			for(int m = 0; m < cnt; m += 2) {
				const bool two = m + 1 < cnt;
				const double *w0 = own + m * OWN_W, *w1 = own + (two ? m + 1 : m) * OWN_W;
				Trial t0 = tr, t1 = tr;
				t0.nx = unid<T>(w0[0]);
				t0.ny = unid<T>(w0[1]);
				t0.nz = unid<T>(w0[2]);
				t1.nx = unid<T>(w1[0]);
				t1.ny = unid<T>(w1[1]);
				t1.nz = unid<T>(w1[2]);
				double P0[12], P1[12];
#pragma unroll
				for(int c = 0; c < 12; c++) {
					P0[c] = w0[NCOMP + c];
					P1[c] = w1[NCOMP + c];
				}
				int b0 = 0, b1 = 0, mb0 = 0, mb1 = 0;
				bool band0 = tr.c2 < 0., band1 = band0;
				if(!band0) {
#pragma unroll
					for(int u = 0; u < CPL; u++) {
						auto lo = [&](double *Oq) { reg_o(u, Oq); };
						if(slot_for(u, m)) {
							int b = 0;
							bonds_new(t0, P0, cxr[u], cyr[u], czr[u], lo, b, band0);
							add(u, b, b0, mb0);
						}
						if(two && slot_for(u, m + 1)) {
							int b = 0;
							bonds_new(t1, P1, cxr[u], cyr[u], czr[u], lo, b, band1);
							add(u, b, b1, mb1);
						}
					}
					each_far([&](double qx, double qy, double qz, auto load_o) {
						bonds_new(t0, P0, qx, qy, qz, load_o, b0, band0);
						if(two) bonds_new(t1, P1, qx, qy, qz, load_o, b1, band1);
					});
				}
				if(__ballot(band0) != 0ull && team_any<T>(band0) && band0) {
					b0 = 0;
					mb0 = 0;
#pragma unroll
					for(int u = 0; u < CPL; u++)
						if(slot_for(u, m)) {
							int b = 0;
							bonds_new_exact(t0, P0, cxr[u], cyr[u], czr[u], [&](double *Oq) { reg_o(u, Oq); }, b);
							add(u, b, b0, mb0);
						}
					each_far([&](double qx, double qy, double qz, auto load_o) {
						bonds_new_exact(t0, P0, qx, qy, qz, load_o, b0);
					});
				}
				if(__ballot(band1) != 0ull && team_any<T>(band1) && band1) {
					b1 = 0;
					mb1 = 0;
#pragma unroll
					for(int u = 0; u < CPL; u++)
						if(slot_for(u, m + 1)) {
							int b = 0;
							bonds_new_exact(t1, P1, cxr[u], cyr[u], czr[u], [&](double *Oq) { reg_o(u, Oq); }, b);
							add(u, b, b1, mb1);
						}
					each_far([&](double qx, double qy, double qz, auto load_o) {
						bonds_new_exact(t1, P1, qx, qy, qz, load_o, b1);
					});
				}
				// (each count is below the team's neighbour count, M < 65536)
				const int b = team_sum<T>(b0 | (b1 << 16));
				if(lane == 0) {
					sbond[m] = b & 0xffff;
					if(two) sbond[m + 1] = b >> 16;
				}
				if(mem_lane) mbits |= ((unsigned) mb0 << m) | (two ? (unsigned) mb1 << (m + 1) : 0u);
			}
			team_sync();
			// The members' fractional coordinates inside the cell, u = (cell
			// coordinate as cell_coord forms it) - (ix, iy, iz), kept in the
			// record's patch slots (the kept form reads no old patches): a
			// trial whose u + displacement stays 1e-9 cells clear of the
			// faces cannot leave the cell (cell_coord's roundings are below
			// 1e-13 cells), so it skips the exact test.
			if(lane < cnt) {
				double *wm = own + lane * OWN_W;
				const double p3[3] = {wm[0], wm[1], wm[2]};
				const double sh[3] = {sh0, sh1, sh2}, il[3] = {tr.iLx, tr.iLy, tr.iLz};
				const int nn[3] = {n0, n1, n2}, ii[3] = {ix, iy, iz};
#pragma unroll
				for(int a = 0; a < 3; a++) {
					double t = (p3[a] - sh[a]) * il[a];
					t = t - __builtin_floor(t);
					wm[NCOMP + a] = t * (double) nn[a] - (double) ii[a];
				}
			}
			team_sync();
		}
		// cells per unit length along each axis (the u step of a displacement)
		const double gnx = (double) n0 * tr.iLx, gny = (double) n1 * tr.iLy, gnz = (double) n2 * tr.iLz;
		int c_acc = 0, d_bonds = 0;
		// this lane's share of a trial's O' and patches (see the trial loop):
		// value lane, and with T < 21 also value lane + T
		static_assert(T >= 16 && CELL_CAP <= 32, "a team of 16 lanes or more");
		const RotShare rsh = rot_share(lane), rsh2 = rot_share(lane + T);

		for(int t = 0; t < ntrial; t++) {
			// the moving member's rank, read from lane t of the team; then one
			// LDS read round: the trial, the member's state and this lane's
			// member
			int k;
			static_assert(T == 16 || T == 32 || T == 64, "teams of 16, 32 or 64 lanes");
			if constexpr(T == 64) k = __builtin_amdgcn_readlane(my_k, t);
			else if constexpr(T == 32) {
				const int k0 = __builtin_amdgcn_readlane(my_k, t), k1 = __builtin_amdgcn_readlane(my_k, t + 32);
				k = koff ? k1 : k0;
			}
			else {
				const int k0 = __builtin_amdgcn_readlane(my_k, t), k1 = __builtin_amdgcn_readlane(my_k, t + 16);
				const int k2 = __builtin_amdgcn_readlane(my_k, t + 32), k3 = __builtin_amdgcn_readlane(my_k, t + 48);
				k = koff < 32 ? (koff ? k1 : k0) : (koff == 32 ? k2 : k3);
			}
			const double *tp = trials + t * TRIAL_W;
			double tv[TRIAL_W];
#pragma unroll
			for(int e = 0; e < TRIAL_W; e++) tv[e] = tp[e];
			// a member slot: the member's current state (every member but the
			// moving one is a neighbour of the trial; a cell of one member
			// reads nothing here)
			const bool is_mem = lane < cnt;
			const bool has_m = is_mem && lane != k;
			if(has_m) {
				const double *wl = own + lane * OWN_W;
				cxr[0] = wl[0];
				cyr[0] = wl[1];
				czr[0] = wl[2];
				if constexpr(REG_ORI)
#pragma unroll
					for(int c = 0; c < 9; c++) orv[0][c] = wl[3 + c];
			}
			// register slot u holds a neighbour of this trial
			auto has_slot = [&](int u) { return u == 0 && is_mem ? has_m : cqr[u] >= 0; };
			const double *me = own + k * OWN_W;
			double ux = 0., uy = 0., uz = 0.;  // (kept form) the member's u
			if constexpr(KEEP) {
				ux = me[NCOMP];
				uy = me[NCOMP + 1];
				uz = me[NCOMP + 2];
			}
			// the rows of the member's O that the lane's share of O' needs
			const double *orow = me + 3 + 3 * rsh.row;
			const double o0 = orow[0], o1 = orow[1], o2 = orow[2];
			double o3 = 0., o4 = 0., o5 = 0.;
			if constexpr(T < 21) {
				const double *orow2 = me + 3 + 3 * rsh2.row;
				o3 = orow2[0];
				o4 = orow2[1];
				o5 = orow2[2];
			}
			double Po[12];  // the old patches (two-sided tests only)
			if constexpr(!KEEP)
#pragma unroll
				for(int c = 0; c < 12; c++) Po[c] = me[NCOMP + c];
			tr.px = unid<T>(me[0]);
			tr.py = unid<T>(me[1]);
			tr.pz = unid<T>(me[2]);
			tr.nx = unid<T>(tr.px + tv[1]);
			tr.ny = unid<T>(tr.py + tv[2]);
			tr.nz = unid<T>(tr.pz + tv[3]);
			// a trial that leaves the cell is rejected
			// (kept form: skipped when u + displacement is clear of the faces)
			double cu0 = 0., cu1 = 0., cu2 = 0.;
			bool deep = false;
			if constexpr(KEEP) {
				cu0 = __builtin_fma(tv[1], gnx, ux);  // u is an estimate: fused is fine
				cu1 = __builtin_fma(tv[2], gny, uy);
				cu2 = __builtin_fma(tv[3], gnz, uz);
				constexpr double CELL_EPS = 1e-9;
				deep = __builtin_fmax(__builtin_fabs(cu0 - 0.5), __builtin_fmax(__builtin_fabs(cu1 - 0.5),
				                                                               __builtin_fabs(cu2 - 0.5))) <
				       0.5 - CELL_EPS;
			}
			if(!deep) {
				// all three axes evaluated (no short circuit)
				const int out = (int) (uni<T>(cell_coord(tr.nx, sh0, tr.iLx, n0)) != ix) |
				                (int) (uni<T>(cell_coord(tr.ny, sh1, tr.iLy, n1)) != iy) |
				                (int) (uni<T>(cell_coord(tr.nz, sh2, tr.iLz, n2)) != iz);
				if(out) continue;
			}

			// phase 1: hard cores at the new position, wavefront-ballot early exit
			// (the register neighbours' separations are kept for phase 2)
			bool ov = false;
			double sx[CPL], sy[CPL], sz[CPL], s2[CPL];
#pragma unroll
			for(int u = 0; u < CPL; u++)
				if(has_slot(u)) {
					s2[u] = sep_new(tr, cxr[u], cyr[u], czr[u], sx[u], sy[u], sz[u]);
					ov |= s2[u] < 1.;
				}
			if(wave_far)
				for(int b = T * CPL - cnt; b < M; b += T)
					if(b + lane < M) {
						const int q = nb_slot(b + lane);
						ov |= overlaps_new(tr, nbld<DD>(F + q), nbld<DD>(F + S + q), nbld<DD>(F + 2 * S + q));
					}
			if(team_any<T>(ov)) continue;

			// O' = O R in the body frame (utils.c:172-184, 204-209) and its
			// patches (utils.h:17-21), one value per lane (two on some lanes of
			// a 16-lane team) in the reference's order: value v < 9 is O'[v],
			// v = 9 + 3a + kk the patch component P'[a][kk] (rot_value).  The
			// values meet in the stage row.
			double rv = 0., rv2 = 0.;
			if(lane < 21) {
				rv = rot_value(rsh, o0, o1, o2, tv);
				stage[lane] = rv;
			}
			if constexpr(T < 21)
				if(lane + T < 21) {
					rv2 = rot_value(rsh2, o3, o4, o5, tv);
					stage[lane + T] = rv2;
				}
			team_sync();
			double Pn[12];
#pragma unroll
			for(int c = 0; c < 12; c++) Pn[c] = stage[9 + c];

			// phase 2: bonds at the old and the new position, over this lane's
			// neighbours: its register slots (a member or static neighbours;
			// KEEP: new position only, the old count is the member's kept one)
			// and the static neighbours beyond them
			int bo, bn, s_new = 0, bnm = 0;
			if constexpr(KEEP) {
				int bns = 0;
				const int bom = has_m ? (int) ((mbits >> k) & 1u) : 0;
				// a register slot's bond: with a member (slot 0 of a member
				// lane) or with a static neighbour
				auto add_slot = [&](int u, int b) {
					if(u == 0) {
						bnm += is_mem ? b : 0;
						bns += is_mem ? 0 : b;
					}
					else bns += b;
				};
				bool band = tr.c2 < 0.;  // no filter: every lane takes the exact form
				if(!band) {
#pragma unroll
					for(int u = 0; u < CPL; u++)
						if(has_slot(u)) {
							int b = 0;
							bonds_new_sep(tr, Pn, sx[u], sy[u], sz[u], s2[u], [&](double *Oq) { reg_o(u, Oq); }, b,
							              band);
							add_slot(u, b);
						}
					each_far([&](double qx, double qy, double qz, auto load_o) {
						bonds_new(tr, Pn, qx, qy, qz, load_o, bns, band);
					});
				}
				if(__ballot(band) != 0ull && team_any<T>(band) && band) {
					bns = 0;
					bnm = 0;
#pragma unroll
					for(int u = 0; u < CPL; u++)
						if(has_slot(u)) {
							int b = 0;
							bonds_new_exact(tr, Pn, cxr[u], cyr[u], czr[u], [&](double *Oq) { reg_o(u, Oq); }, b);
							add_slot(u, b);
						}
					each_far([&](double qx, double qy, double qz, auto load_o) {
						bonds_new_exact(tr, Pn, qx, qy, qz, load_o, bns);
					});
				}
				// three counts below 1024 in one reduction
				const int packed = team_sum<T>(bns | (bnm << 10) | (bom << 20));
				s_new = packed & 1023;
				bo = sbond[k] + (packed >> 20);
				bn = s_new + ((packed >> 10) & 1023);
			}
			else {
				bo = 0;
				bn = 0;
				bool band = tr.c2 < 0.;
				auto each_neighbour = [&](auto &&f) {
#pragma unroll
					for(int u = 0; u < CPL; u++)
						if(has_slot(u)) f(cxr[u], cyr[u], czr[u], [&](double *Oq) { reg_o(u, Oq); });
					each_far(f);
				};
				if(!band)
					each_neighbour([&](double qx, double qy, double qz, auto load_o) {
						bonds_old_new(tr, Po, Pn, qx, qy, qz, load_o, bo, bn, band);
					});
				if(__ballot(band) != 0ull && team_any<T>(band) && band) {
					bo = 0;
					bn = 0;
					each_neighbour([&](double qx, double qy, double qz, auto load_o) {
						bonds_old_new_exact(tr, Po, Pn, qx, qy, qz, load_o, bo, bn);
					});
				}
				bo = team_sum_small<T>(bo);
				bn = team_sum_small<T>(bn);
			}
			const int db = bn - bo;  // E = -bonds, dE = -db
			const double dE = (double) (-db);
			// MC.c:332: accept if dE < 0 or a uniform draw u < exp(-dE/T); dE
			// is an integer: below ACC_TAB the trial's acceptance bits hold
			// the comparison with the replica's exp(-dE/T)
			const int idE = -db;
			const unsigned am = (unsigned) tv[0];
			const bool acc = idE < 0 ||
			                 (idE == 0 ? tv[13] < 1.
			                           : (idE < ACC_TAB ? ((am >> idE) & 1u) != 0u : tv[13] < exp(-dE / temp)));
			if(acc) {
				c_acc++;
				d_bonds += db;
				// member k takes the new position, orientation and patches
				// (O' and P' are contiguous in its record, lane l's at 3 + l)
				double *wk = own + k * OWN_W;
				// (the kept form keeps u in the patch slots instead of P')
				constexpr int NW = KEEP ? 9 : 21;
				if(lane < NW) wk[3 + lane] = rv;
				if constexpr(T < NW)
					if(lane + T < NW) wk[3 + lane + T] = rv2;
				// (orientations in LDS) and into member k's register slot
				if constexpr(!REG_ORI)
					if(lane < 9) ori[lane * T + k] = rv;
				if(lane == 0) {
					wk[0] = tr.nx;
					wk[1] = tr.ny;
					wk[2] = tr.nz;
					if constexpr(KEEP) {
						wk[NCOMP] = cu0;
						wk[NCOMP + 1] = cu1;
						wk[NCOMP + 2] = cu2;
					}
					if constexpr(KEEP) sbond[k] = s_new;
				}
				if constexpr(KEEP) {
					// member k's bonds with the others are now the new ones
					const unsigned long long nb_bits =
					    (__ballot(bnm != 0) >> koff) & (T >= 32 ? 0xffffffffull : ((1ull << T) - 1ull));
					if(lane < cnt) mbits = lane == k ? (unsigned) nb_bits : ((mbits & ~(1u << k)) | ((unsigned) bnm << k));
				}
			}
			team_sync();
		}
		// 6. write the members back, ranked (the ranking permutes the cell's slots)
		if(c_acc > 0 || cnt > 1) {
			for(int e = lane; e < cnt * NCOMP; e += T) {
				int m = e / NCOMP, c = e - m * NCOMP;
				if constexpr(DD)
					__hip_atomic_store(F + c * S + s0 + m, own[m * OWN_W + c], __ATOMIC_RELAXED,
					                   __HIP_MEMORY_SCOPE_AGENT);
				else F[c * S + s0 + m] = own[m * OWN_W + c];
			}
			if(lane < cnt) {
				int rank = 0;
				for(int m = 0; m < cnt; m++) rank += own_pid[m] < my_pid;
				PID[s0 + rank] = my_pid;
			}
		}
		if constexpr(DD) {
			// the members' stores have completed (the wavefront's memory
			// counter is zero) before the cell is published as done
			__atomic_signal_fence(__ATOMIC_SEQ_CST);
			__builtin_amdgcn_s_waitcnt(0);
			__atomic_signal_fence(__ATOMIC_SEQ_CST);
			if(lane == 0)
				__hip_atomic_store(d.done + rp.cell_off + cloc, mark, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		}
		if(lane == 0) {
			const int slot = gi & (NSLOT - 1);
			atomicAdd(ctr_at(d.ctr, r, C_TRY_ROTO, slot), (unsigned long long) ntrial);
			if(c_acc) {
				atomicAdd(ctr_at(d.ctr, r, C_ACC_ROTO, slot), (unsigned long long) c_acc);
				atomicAdd(ctr_at(d.ctr, r, C_DBONDS, slot), (unsigned long long) (long long) d_bonds);
			}
		}
	}
}

// One colour pass per launch, one list entry per team.  The grid is sized
// for the largest list; workgroups reach the 8 XCDs round-robin, and the
// list is in cell order, so XCD x takes the x-th eighth of the entries
// actually listed (a slab of the box, whose shared neighbour reads its L2
// then serves); the surplus workgroups of each XCD leave at once.
template <int T, int CPL, bool KEEP, bool ONE, int CAP = CELL_CAP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_pass(PassArgs a) {
	extern __shared__ double lds[];
	const int lane = threadIdx.x & (T - 1);
	const int team = threadIdx.x / T;
	const int tpb = blockDim.x / T;
	const int total = a.d.pfx[a.pass * (a.d.R + 1) + a.d.R];
	const int per_xcd = ((total + tpb - 1) / tpb + 7) >> 3;  // workgroups with work on each XCD
	const int i = blockIdx.x >> 3;
	if(i >= per_xcd) return;
	const int gi = uni<T>(((blockIdx.x & 7) * per_xcd + i) * tpb + team);
	if(gi >= total) return;
	// 16-lane teams with two slots per lane keep the reachable neighbours only
	pass_cell<T, CPL, KEEP, ONE, false, T == 16 && CPL == 2 && KEEP, CAP>(a.d, a.sweep, a.pass, gi, lane,
	                                                                       lds + team * team_lds<T, CAP>(), 0u);
}

// The cells the small-cell pass deferred (more than SMALL_CAP members), by
// the general 32-lane form: teams take the listed entries in turn.
template <bool KEEP, bool ONE>
__global__ __launch_bounds__(128) void k_pass_ovf(PassArgs a) {
	extern __shared__ double lds[];
	constexpr int T = 32;
	const int lane = threadIdx.x & (T - 1);
	const int team = threadIdx.x / T;
	const int tpb = blockDim.x / T;
	const int n = a.d.ovf_n[a.pass];
	for(int e = blockIdx.x * tpb + team; e < n; e += gridDim.x * tpb)
		pass_cell<T, 1, KEEP, ONE, false>(a.d, a.sweep, a.pass, a.d.ovf[e], lane, lds + team * team_lds<T>(), 0u);
}

// The 8 colour passes of a sweep in one launch, driven by dependencies
// instead of kernel boundaries.  The sweep's work is one queue of tickets:
// the list of pass 0, then pass 1, ..., each padded to whole wavefronts so
// that the teams of one wavefront always work on one pass.  Each wavefront
// takes the next tickets when it starts; each team then runs its cell as the
// per-pass kernel does, after waiting for its neighbours of earlier passes
// (wait_neighbours), and publishes its own cell.  A cell's neighbours of
// earlier passes hold earlier tickets, taken by wavefronts that started
// before it and are running or done, so the work always progresses whatever
// is resident; the waits are bounded besides.  The grid covers the largest
// possible lists; wavefronts whose tickets lie past the sweep's end leave at
// once.  Results equal the per-pass launches' bit for bit: each cell sees its
// neighbourhood as the earlier passes left it, and its trials do not depend
// on timing.
template <int T, int CPL, bool KEEP, bool ONE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_pass_dd(PassArgs a) {
	extern __shared__ double lds[];
	const DevState &d = a.d;
	constexpr int TPW = 64 / T;  // teams per wavefront
	const int lane = threadIdx.x & (T - 1);
	const int team = threadIdx.x / T;
	int t0 = 0;
	if((threadIdx.x & 63) == 0) t0 = atomicAdd(d.ddq, TPW);
	const int t = __builtin_amdgcn_readfirstlane(t0) + (threadIdx.x & 63) / T;
	int p = 0, base = 0, n = 0;
	for(; p < 8; p++) {
		n = d.pfx[p * (d.R + 1) + d.R];
		const int padded = (n + TPW - 1) / TPW * TPW;
		if(t < base + padded) break;
		base += padded;
	}
	if(p == 8 || t - base >= n) return;
	pass_cell<T, CPL, KEEP, ONE, true>(d, a.sweep, p, t - base, lane, lds + team * team_lds<T>(), a.mark);
}

// Grid-wide barrier of a cooperative launch (every workgroup resident).
// Arrivals count up in bar[0]; the last one resets it and advances the
// generation bar[1], which the others wait for.  The agent-scope fences
// publish each workgroup's writes before its arrival and make the others'
// visible after the wait.  The wait is bounded: past BAR_SPIN_MAX polls the
// workgroup reports ERR_BARRIER and goes on, so the grid always drains.
__device__ __forceinline__ void grid_barrier(unsigned *bar, int *err) {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
	__syncthreads();
	if(threadIdx.x == 0) {
		const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		const unsigned arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		if(arrived == gridDim.x - 1) {
			__hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			__hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
		}
		else {
			int spins = 0;
			while(__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
				if(++spins > BAR_SPIN_MAX) {
					atomicOr(err, (int) ERR_BARRIER);
					break;
				}
				__builtin_amdgcn_s_sleep(2);
			}
		}
	}
	__syncthreads();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// The 8 colour passes of a sweep in one cooperative launch, separated by
// grid barriers.  Team g takes list entries g, g + teams, ... of each pass.
template <int T, int CPL, bool KEEP>
__global__ __launch_bounds__(256) void k_sweep_passes(PassArgs a) {
	extern __shared__ double lds[];
	const DevState &d = a.d;
	const int lane = threadIdx.x & (T - 1);
	const int team = threadIdx.x / T;
	const int tpb = blockDim.x / T;
	const int nteams = gridDim.x * tpb;
	const int g0 = uni<T>(blockIdx.x * tpb + team);
	double *own = lds + team * team_lds<T>();
	for(int k = 0; k < 8; k++) {
		const int total = uni<T>(d.pfx[k * (d.R + 1) + d.R]);
		for(int gi = g0; gi < total; gi += nteams) pass_cell<T, CPL, KEEP, false, false>(d, a.sweep, k, gi, lane, own, 0u);
		if(k < 7) grid_barrier(d.bar, d.err);
	}
}

template <int T, int CPL, bool KEEP, bool ONE>
static void pass_dd_T(const DevState &d, uint64_t sweep, unsigned mark, int block, hipStream_t s) {
	const int tpb = block / T;
	// every pass's largest list, one team per entry (wavefronts past the
	// sweep's tickets leave at once)
	const unsigned nb = (unsigned) std::max(1, 8 * ((d.total_act + tpb - 1) / tpb));
	size_t lds = (size_t) tpb * team_lds<T>() * sizeof(double);
	PassArgs a{d, sweep, 0, mark};
	hipLaunchKernelGGL((k_pass_dd<T, CPL, KEEP, ONE>), dim3(nb), dim3(block), lds, s, a);
}

template <int T, bool KEEP>
static void pass_dd_TK(const DevState &d, uint64_t sweep, unsigned mark, int block, hipStream_t s) {
	if(d.R == 1) pass_dd_T<T, 1, KEEP, true>(d, sweep, mark, block, s);
	else pass_dd_T<T, 1, KEEP, false>(d, sweep, mark, block, s);
}

void launch_passes_dd(const DevState &d, uint64_t sweep, unsigned mark, int T, int block, hipStream_t s) {
	const bool keep = d.trials_per_cell >= KEEP_MIN_TRIALS;
	if(T == 32) {
		if(keep) pass_dd_TK<32, true>(d, sweep, mark, block, s);
		else pass_dd_TK<32, false>(d, sweep, mark, block, s);
	}
	else {
		if(keep) pass_dd_TK<64, true>(d, sweep, mark, block, s);
		else pass_dd_TK<64, false>(d, sweep, mark, block, s);
	}
}

template <int T, int CPL, bool KEEP, bool ONE, int CAP = CELL_CAP>
static void pass_T(const DevState &d, uint64_t sweep, int pass, int block, hipStream_t s) {
	const int tpb = block / T;
	int nb = (d.total_act + tpb - 1) / tpb;
	nb = (nb + 7) & ~7;  // 8 XCDs x the most workgroups one XCD may need
	size_t lds = (size_t) tpb * team_lds<T, CAP>() * sizeof(double);
	PassArgs a{d, sweep, pass, 0u};
	hipLaunchKernelGGL((k_pass<T, CPL, KEEP, ONE, CAP>), dim3(nb), dim3(block), lds, s, a);
	if constexpr(CAP < CELL_CAP) {
		// the deferred cells (usually none: the workgroups leave at once)
		constexpr int OB = 128;
		hipLaunchKernelGGL((k_pass_ovf<KEEP, ONE>), dim3(256), dim3(OB), (size_t) (OB / 32) * team_lds<32>() * sizeof(double),
		                   s, a);
	}
}

int small_cell_cap() { return SMALL_CAP; }

size_t pass_lds_bytes(int T, int block) {
	return (size_t) (block / T) * (T == 16 ? team_lds<16>() : (T == 32 ? team_lds<32>() : team_lds<64>())) *
	       sizeof(double);
}

// The form is chosen per context: 32 or 64 lanes per cell, kept static
// bonds or not, and one replica (its constants then load to scalar registers
// for every team) or several.
template <int T, bool KEEP>
static void pass_TK(const DevState &d, uint64_t sweep, int pass, int block, hipStream_t s) {
	if constexpr(T == 16) {
		// dilute systems (fewer than 16 candidates in a cell's 27 cells on
		// average): one register neighbour per lane, the rare rest from memory
		const double cand = 27. * (double) d.Ntot / (double) std::max(d.C_total, 1);
		if(cand < 16.) {
			if(d.R == 1) pass_T<16, 1, KEEP, true, SMALL_CAP>(d, sweep, pass, block, s);
			else pass_T<16, 1, KEEP, false, SMALL_CAP>(d, sweep, pass, block, s);
			return;
		}
	}
	constexpr int CPL = T < 32 ? 32 / T : 1;  // register neighbours per lane
	if(d.R == 1) pass_T<T, CPL, KEEP, true>(d, sweep, pass, block, s);
	else pass_T<T, CPL, KEEP, false>(d, sweep, pass, block, s);
}

void launch_pass(const DevState &d, uint64_t sweep, int pass, int T, int block, hipStream_t s) {
	const bool keep = d.trials_per_cell >= KEEP_MIN_TRIALS;
	if(T == 16) {
		if(keep) pass_TK<16, true>(d, sweep, pass, block, s);
		else pass_TK<16, false>(d, sweep, pass, block, s);
	}
	else if(T == 32) {
		if(keep) pass_TK<32, true>(d, sweep, pass, block, s);
		else pass_TK<32, false>(d, sweep, pass, block, s);
	}
	else {
		if(keep) pass_TK<64, true>(d, sweep, pass, block, s);
		else pass_TK<64, false>(d, sweep, pass, block, s);
	}
}

template <int T, int CPL, bool KEEP>
static int sweep_blocks_T(int block, int cus) {
	int per_cu = 0;
	if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *) k_sweep_passes<T, CPL, KEEP>, block,
	                                                pass_lds_bytes(T, block)) != hipSuccess)
		return 0;
	return per_cu * cus;
}

// co-resident workgroups of either form (both must fit: the form is chosen
// per context at launch)
int sweep_blocks_max(int T, int block, int cus) {
	if(T == 32) return std::min(sweep_blocks_T<32, 1, false>(block, cus), sweep_blocks_T<32, 1, true>(block, cus));
	return std::min(sweep_blocks_T<64, 1, false>(block, cus), sweep_blocks_T<64, 1, true>(block, cus));
}

template <int T, int CPL, bool KEEP>
static hipError_t sweep_T(const DevState &d, uint64_t sweep, int block, int blocks, hipStream_t s) {
	PassArgs a{d, sweep, 0, 0u};
	void *args[] = {&a};
	return hipLaunchCooperativeKernel((const void *) k_sweep_passes<T, CPL, KEEP>, dim3(blocks), dim3(block), args,
	                                  (unsigned) pass_lds_bytes(T, block), s);
}

hipError_t launch_sweep_passes(const DevState &d, uint64_t sweep, int T, int block, int blocks, hipStream_t s) {
	const bool keep = d.trials_per_cell >= KEEP_MIN_TRIALS;
	if(T == 32) return keep ? sweep_T<32, 1, true>(d, sweep, block, blocks, s) : sweep_T<32, 1, false>(d, sweep, block, blocks, s);
	return keep ? sweep_T<64, 1, true>(d, sweep, block, blocks, s) : sweep_T<64, 1, false>(d, sweep, block, blocks, s);
}

// ------------------------------------------------------------------------
// energy: one team per slot, MC_energy semantics (MC.c:272-309): bonds of
// the particle with every neighbour, 0 and an overlap flag on any overlap.

template <int T>
struct NbLayout {
	static constexpr int S = (T >= 54) ? 2 : 1;
	static constexpr int STEP = T / S;
	static constexpr int PER = (27 + STEP - 1) / STEP;
};

// trial = 1: the energy of the replicas' box trials.  A trial that rebuilt
// its lists reads the spare buffer on the trial grid; a fast trial reads the
// current buffer through the current lists (membership by the list grid,
// computed from the unscaled positions exactly as the lists were built) and
// scales every position as the trial does before it measures separations.
// Once one team of a trial has met an overlap the trial is rejected
// whatever the bonds are: the replica's later teams stop at once.
// Workgroups reach the XCDs round-robin: each XCD takes a contiguous range of
// slots (neighbouring cells share its L2).
// The bond decision of a pair of particles closer than the bond range but
// not overlapping (MC_would_interact's angular part, MC.c:248-262), seen from
// p with separation (dx, dy, dz) towards q and p's patches P: the filtered
// test, and the reference's arithmetic only inside the filter's band or
// without the filter.  Seen from q the decision is the same (every step is
// the exact negation of p's).
template <class LoadO>
__device__ __forceinline__ bool pair_bonded(double dx, double dy, double dz, double d2, const double *P, double c2,
                                            double cosmax, LoadO load_q) {
	bool unsure = c2 < 0.;
	bool bonded = false;
	if(!unsure) {
		const KfBand bd = kf_band(c2, d2);
		const unsigned fp = kf_fast_p(dx, dy, dz, P, bd);
		unsure = (fp >> 1) != 0u;
		if(fp & 1u) {
			double Oq[9];
			load_q(Oq);
			const unsigned fq = kf_fast_q1(dx, dy, dz, kf_band_q(bd), Oq);
			unsure |= (fq >> 1) != 0u;
			if(!unsure) bonded = (fq & 1u) != 0u;
		}
	}
	if(unsure) {
		kf_unit(dx, dy, dz, d2);
		bonded = false;
		if(kf_any_p(dx, dy, dz, P, cosmax)) {
			double Oq[9];
			load_q(Oq);
			bonded = kf_any_q(dx, dy, dz, Oq, cosmax);
		}
	}
	return bonded;
}

template <int T>
__global__ __launch_bounds__(256) void k_energy(DevState d, int trial, int per_particle) {
	using NB = NbLayout<T>;
	const int lane = threadIdx.x & (T - 1);
	const int team = threadIdx.x / T;
	const int j = xcd_remap(blockIdx.x, gridDim.x) * (blockDim.x / T) + team;
	if(j >= d.Ntot) return;
	const int r = d.prep[j];
	const Rep &rp = d.reps[r];
	if(j - rp.p_off >= rp.np) return;  // a free slot (the live particles come first after a rebuild)
	const bool tr = trial && rp.trial_active;
	if(trial && !tr) return;
	if(tr && (rp.trial_rec || rp.trial_pl)) return;  // k_energy_rec / k_energy_pairs take this trial
	if(tr && __hip_atomic_load(&d.vflag[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
	const bool fast = tr && rp.trial_fast;
	const Grid &G = tr ? (fast ? rp.lg : rp.tg) : rp.g;  // the grid of the lists' membership
	const Grid &B = tr ? rp.tg : rp.g;                    // the box of the separations
	const double s0 = fast ? rp.trial_s[0] : 1., s1 = fast ? rp.trial_s[1] : 1., s2 = fast ? rp.trial_s[2] : 1.;
	const int b = (tr && !fast) ? (rp.sel ^ 1) : rp.sel;
	const double *F = pick_buf(d.f, b);
	const int S = d.stride;
	const double ux = F[j], uy = F[S + j], uz = F[2 * S + j];
	double O[9];
#pragma unroll
	for(int c = 0; c < 9; c++) O[c] = F[(3 + c) * S + j];
	double P[4][3];
	kf_patches(O, P);
	const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
	const int ix = cell_coord(ux, G.shift[0], G.inv_box[0], n0);
	const int iy = cell_coord(uy, G.shift[1], G.inv_box[1], n1);
	const int iz = cell_coord(uz, G.shift[2], G.inv_box[2], n2);
	// (x * 1 is x: unscaled replicas are unchanged)
	const double px = ux * s0, py = uy * s1, pz = uz * s2;
	const double Lx = B.box[0], Ly = B.box[1], Lz = B.box[2];
	const double iLx = B.inv_box[0], iLy = B.inv_box[1], iLz = B.inv_box[2];
	const double c2 = rp.cosmax >= KF_FILTER_MIN_COS ? rp.cosmax * rp.cosmax : -1.;
	int nbond = 0;
	bool ov = false;
#pragma unroll
	for(int u = 0; u < NB::PER; u++) {
		int nb = lane / NB::S + u * NB::STEP;
		if(nb >= 27) continue;
		int cx = wrapi(ix + nb / 9 - 1, n0), cy = wrapi(iy + (nb / 3) % 3 - 1, n1), cz = wrapi(iz + nb % 3 - 1, n2);
		int cc = rp.cell_off + (cx * n1 + cy) * n2 + cz;
		int s = d.start[cc], e = d.start[cc + 1];
		for(int q = s + lane % NB::S; q < e; q += NB::S) {
			if(q == j) continue;
			double dx = kf_min_image(F[q] * s0 - px, Lx, iLx);
			double dy = kf_min_image(F[S + q] * s1 - py, Ly, iLy);
			double dz = kf_min_image(F[2 * S + q] * s2 - pz, Lz, iLz);
			double d2 = dx * dx + dy * dy + dz * dz;
			if(d2 < 1.) ov = true;
			else if(d2 < rp.rc2)
				nbond += (int) pair_bonded(dx, dy, dz, d2, &P[0][0], c2, rp.cosmax, [&](double *Oq) {
#pragma unroll
					for(int c = 0; c < 9; c++) Oq[c] = F[(3 + c) * S + q];
				});
		}
	}
	nbond = team_sum<T>(nbond);
	const bool any = team_any<T>(ov);
	if(lane == 0) {
		if(tr && any) __hip_atomic_store(&d.vflag[r], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		int res = any ? 0 : nbond;
		if(per_particle) {
			int i = rp.p_off + pick_buf(d.pid, b)[j];
			d.out_bonds[i] = res;
			d.out_ovl[i] = any ? 1 : 0;
		}
		int slot = j & (NSLOT - 1);
		if(res) atomicAdd(ctr_at(d.ctr, r, C_EBONDS, slot), (unsigned long long) res);
		if(any) atomicAdd(ctr_at(d.ctr, r, C_OVL, slot), 1ull);
	}
}

// Bonded partners of every particle (the pairs output_print_bonds lists,
// output.c:213-268, with MC_interact's test): slot j's partners' particle
// indices at partners[(p_off + pid) * max_per + k], k < max_per, and their
// number at count[p_off + pid] (which may exceed max_per).  One thread per
// slot; an output path, not a hot one.
__global__ void k_bond_list(DevState d, int max_per, int *count, int *partners) {
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= d.Ntot) return;
	const int r = d.prep[j];
	const Rep &rp = d.reps[r];
	if(j - rp.p_off >= rp.np) return;
	const Grid &G = rp.g;
	const double *F = pick_buf(d.f, rp.sel);
	const int *PID = pick_buf(d.pid, rp.sel);
	const int S = d.stride;
	const double px = F[j], py = F[S + j], pz = F[2 * S + j];
	double O[9];
	for(int c = 0; c < 9; c++) O[c] = F[(3 + c) * S + j];
	double P[4][3];
	kf_patches(O, P);
	const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
	const int ix = cell_coord(px, G.shift[0], G.inv_box[0], n0);
	const int iy = cell_coord(py, G.shift[1], G.inv_box[1], n1);
	const int iz = cell_coord(pz, G.shift[2], G.inv_box[2], n2);
	const size_t me = (size_t) rp.p_off + PID[j];
	int nb = 0;
	for(int c = 0; c < 27; c++) {
		const int cx = wrapi(ix + c / 9 - 1, n0), cy = wrapi(iy + (c / 3) % 3 - 1, n1), cz = wrapi(iz + c % 3 - 1, n2);
		const int cc = rp.cell_off + (cx * n1 + cy) * n2 + cz;
		for(int q = d.start[cc]; q < d.start[cc + 1]; q++) {
			if(q == j) continue;
			double dx = kf_min_image(F[q] - px, G.box[0], G.inv_box[0]);
			double dy = kf_min_image(F[S + q] - py, G.box[1], G.inv_box[1]);
			double dz = kf_min_image(F[2 * S + q] - pz, G.box[2], G.inv_box[2]);
			const double d2 = dx * dx + dy * dy + dz * dz;
			if(d2 < 1. || !(d2 < rp.rc2)) continue;
			kf_unit(dx, dy, dz, d2);
			if(!kf_any_p(dx, dy, dz, &P[0][0], rp.cosmax)) continue;
			double Oq[9];
			for(int e = 0; e < 9; e++) Oq[e] = F[(3 + e) * S + q];
			if(!kf_any_q(dx, dy, dz, Oq, rp.cosmax)) continue;
			if(nb < max_per) partners[me * max_per + nb] = PID[q];
			nb++;
		}
	}
	count[me] = nb;
}

void launch_bond_list(const DevState &d, int max_per, int *count, int *partners, hipStream_t s) {
	if(d.Ntot > 0)
		hipLaunchKernelGGL(k_bond_list, dim3((d.Ntot + 255) / 256), dim3(256), 0, s, d, max_per, count, partners);
}

void launch_energy_rows_zero(const DevState &d, hipStream_t s) {
	hipLaunchKernelGGL(k_zero_energy_rows, dim3((d.R * 2 * NSLOT + 255) / 256), dim3(256), 0, s, d.ctr, d.R,
	                   (const int *) (d.vslow ? d.vslow + 1 : nullptr));
}

// Lanes per particle: 32 where the 27 cells hold many candidates, 8 in dilute
// systems (fewer than 8 candidates on average over the cell capacity), where
// most of a 32-lane team would idle.  The result does not depend on it.
void launch_energy(const DevState &d, int trial, int per_particle, hipStream_t s) {
	launch_energy_rows_zero(d, s);
	if(d.Ntot <= 0) return;
	const int block = 256;
	const double cand = 27. * (double) d.Ntot / (double) (d.C_total > 0 ? d.C_total : 1);
	// a multiple of 8 workgroups: one contiguous range of slots per XCD
	auto grid = [&](int per) { return dim3((unsigned) ((((long long) d.Ntot + per - 1) / per + 7) & ~7LL)); };
	if(cand < 8.) hipLaunchKernelGGL(k_energy<8>, grid(block / 8), dim3(block), 0, s, d, trial, per_particle);
	else hipLaunchKernelGGL(k_energy<32>, grid(block / 32), dim3(block), 0, s, d, trial, per_particle);
}

// The energy of box trials on the current lists, from the sweep's cell
// records instead of per particle: listed cell i (of the 8 passes' lists one
// after another) pairs its members with each other once and with the
// members of its neighbours in the 13 directions m > 13 of the record (m =
// 9 (dx + 1) + 3 (dy + 1) + dz + 1), so every pair of particles in adjacent
// cells is seen exactly once.  The separation is formed from the cell's
// member as k_energy forms it from either particle (scaled positions, the
// trial box) and pair_bonded decides the bond, so a bonded pair adds 2 to
// the bond rows (the per-particle sum counts it twice) and an overlapping
// pair adds 1 to the overlap row: k_vol_decide reads the same totals.  Per
// listed cell this reads its record once, where k_energy reads 27 cell ranges
// per particle; on dilute systems most of them are empty.
// TT lanes per listed cell: 8 on dense systems; 1 on dilute ones, where a
// listed cell holds about one particle and one candidate, so that the lanes
// stream through the records instead of idling in teams.
template <int TT>
__global__ __launch_bounds__(256) void k_energy_rec(DevState d) {
	const int lane = threadIdx.x & (TT - 1);
	const int teams = gridDim.x * (blockDim.x / TT);
	int lens[8];
	int total = 0;
#pragma unroll
	for(int p = 0; p < 8; p++) {
		lens[p] = d.pfx[p * (d.R + 1) + d.R];
		total += lens[p];
	}
	for(int i0 = xcd_remap(blockIdx.x, gridDim.x) * (blockDim.x / TT) + threadIdx.x / TT; i0 < total; i0 += teams) {
		int p = 0, gi = i0;
		while(gi >= lens[p]) gi -= lens[p++];
		const int *rec = d.rec + ((size_t) p * d.total_act + gi) * REC_W;
		const int r = rec[62];
		const Rep &rp = d.reps[r];
		if(!rp.trial_active || !rp.trial_rec) continue;
		// (the early exit may read a stale flag: it only saves work)
		if(__hip_atomic_load(&d.vflag[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) continue;
		const int s0 = rec[59], cnt = rec[60];
		const int m0 = rec[32 + 13], m2 = rec[58] - m0;  // candidates in the 13 directions
		const double *F = pick_buf(d.f, rec[63]);
		const int S = d.stride;
		const double t0 = rp.trial_s[0], t1 = rp.trial_s[1], t2 = rp.trial_s[2];
		const double Lx = rp.tg.box[0], Ly = rp.tg.box[1], Lz = rp.tg.box[2];
		const double iLx = rp.tg.inv_box[0], iLy = rp.tg.inv_box[1], iLz = rp.tg.inv_box[2];
		const double c2 = rp.cosmax >= KF_FILTER_MIN_COS ? rp.cosmax * rp.cosmax : -1.;
		int nbond = 0, novl = 0;
		const int work = cnt * m2 + cnt * cnt;
		for(int e = lane; e < work; e += TT) {
			int j, q;
			if(e < cnt * m2) {
				j = s0 + e / m2;
				const int k = e % m2 + m0;  // prefix position among the record's 26 neighbours
				int nb = 13;
				while(nb < 25 && rec[32 + nb + 1] <= k) nb++;
				q = rec[nb] + (k - rec[32 + nb]);
			}
			else {
				const int a = (e - cnt * m2) / cnt, b2 = (e - cnt * m2) % cnt;
				if(a >= b2) continue;
				j = s0 + a;
				q = s0 + b2;
			}
			const double px = F[j] * t0, py = F[S + j] * t1, pz = F[2 * S + j] * t2;
			const double dx = kf_min_image(F[q] * t0 - px, Lx, iLx);
			const double dy = kf_min_image(F[S + q] * t1 - py, Ly, iLy);
			const double dz = kf_min_image(F[2 * S + q] * t2 - pz, Lz, iLz);
			const double d2 = dx * dx + dy * dy + dz * dz;
			if(d2 < 1.) novl++;
			else if(d2 < rp.rc2) {
				double O[9], P[4][3];
#pragma unroll
				for(int c = 0; c < 9; c++) O[c] = F[(3 + c) * S + j];
				kf_patches(O, P);
				nbond += (int) pair_bonded(dx, dy, dz, d2, &P[0][0], c2, rp.cosmax, [&](double *Oq) {
#pragma unroll
					for(int c = 0; c < 9; c++) Oq[c] = F[(3 + c) * S + q];
				});
			}
		}
		if constexpr(TT > 1) {
			nbond = team_sum<TT>(nbond);
			novl = team_sum<TT>(novl);
		}
		if(lane == 0) {
			const int slot = i0 & (NSLOT - 1);
			if(novl) {
				__hip_atomic_store(&d.vflag[r], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				atomicAdd(ctr_at(d.ctr, r, C_OVL, slot), (unsigned long long) novl);
			}
			else if(nbond)
				atomicAdd(ctr_at(d.ctr, r, C_EBONDS, slot), 2ull * (unsigned long long) nbond);
		}
	}
}

void launch_energy_rec(const DevState &d, int cus, hipStream_t s) {
	// a grid-stride loop over the listed cells: 8 workgroups per CU (a
	// multiple of 8: one contiguous range of cells per XCD); one lane per
	// cell when the 27 cells around a particle hold fewer than 8 candidates
	const int blocks = std::max(8, cus * 8) & ~7;
	const double cand = 27. * (double) d.Ntot / (double) (d.C_total > 0 ? d.C_total : 1);
	if(cand < 8.) hipLaunchKernelGGL(k_energy_rec<1>, dim3((unsigned) blocks), dim3(256), 0, s, d);
	else hipLaunchKernelGGL(k_energy_rec<8>, dim3((unsigned) blocks), dim3(256), 0, s, d);
}

// ------------------------------------------------------------------------
// Close pairs of the box trials.  At the start of a sweep's box trials the
// records give every pair of particles in adjacent cells once, as
// k_energy_rec enumerates them; of each replica that makes box trials, the
// pairs closer than PAIR_LIST_FACTOR bond ranges (minimum image in the
// current box) are listed.  A trial scales every separation by the axis
// factors accepted since the list was built (pl_scale) times its own.  While
// the smallest of those products s keeps s x PAIR_LIST_FACTOR above 1 (with a
// margin for the roundings of the list's test), an unlisted pair of adjacent
// cells stays beyond the bond range, and a fast trial already holds every
// pair within the bond range in adjacent cells: the trial's energy is the sum
// over the listed pairs, pair by pair as k_energy_rec takes it.

// The pair of work item e of a record (k_energy_rec's enumeration): the
// cell's member j with candidate k of the 13 directions m > 13, or two
// members; false for the half of the member pairs that is skipped.
__device__ __forceinline__ bool rec_pair(const int *rec, int s0, int cnt, int m0, int m2, int e, int &j, int &q) {
	if(e < cnt * m2) {
		j = s0 + e / m2;
		const int k = e % m2 + m0;
		int nb = 13;
		while(nb < 25 && rec[32 + nb + 1] <= k) nb++;
		q = rec[nb] + (k - rec[32 + nb]);
		return true;
	}
	const int a = (e - cnt * m2) / cnt, b2 = (e - cnt * m2) % cnt;
	j = s0 + a;
	q = s0 + b2;
	return a < b2;
}

// Which replicas list their pairs this sweep (box trials on current lists
// whose records were written by this sweep's rebuild), and an empty list.
__global__ void k_pairs_reset(DevState d) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if(r == 0) {
		d.pl_n[0] = 0;
		d.pl_n[1] = 0;
	}
	if(r >= d.R) return;
	Rep &rp = d.reps[r];
	rp.pl_ok = ((rp.ensemble == 5 || rp.lx_move) && rp.rec_ok && rp.lists_ok) ? 1 : 0;
	rp.pl_scale[0] = rp.pl_scale[1] = rp.pl_scale[2] = 1.;
}

// One lane per listed cell, the wavefront's lanes in step: each counts its
// close pairs, one atomic per wavefront reserves their places (in no fixed
// order: the trial energies are integer sums), then each writes them.  A
// list that would pass pl_cap is marked as overflowed; the trials then take
// their energy from the records.
__global__ __launch_bounds__(256) void k_pairs_build(DevState d) {
	const int lane = threadIdx.x & 63;
	const int wave = (xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + (int) threadIdx.x) >> 6;
	const int waves = gridDim.x * (blockDim.x >> 6);
	int lens[8];
	int total = 0;
#pragma unroll
	for(int p = 0; p < 8; p++) {
		lens[p] = d.pfx[p * (d.R + 1) + d.R];
		total += lens[p];
	}
	for(int w0 = wave * 64; w0 < total; w0 += waves * 64) {
		const int i0 = w0 + lane;
		const int *rec = d.rec;
		bool on = false;
		int s0 = 0, cnt = 0, m0 = 0, m2 = 0;
		const double *F = d.f[0];
		double L[3] = {1., 1., 1.}, iL[3] = {1., 1., 1.}, rl2 = 0.;
		if(i0 < total) {
			int p = 0, gi = i0;
			while(gi >= lens[p]) gi -= lens[p++];
			rec = d.rec + ((size_t) p * d.total_act + gi) * REC_W;
			const Rep &rp = d.reps[rec[62]];
			on = rp.pl_ok != 0;
			if(on) {
				s0 = rec[59];
				cnt = rec[60];
				m0 = rec[32 + 13];
				m2 = rec[58] - m0;
				F = pick_buf(d.f, rec[63]);
				for(int a = 0; a < 3; a++) {
					L[a] = rp.g.box[a];
					iL[a] = rp.g.inv_box[a];
				}
				rl2 = rp.rc2 * (PAIR_LIST_FACTOR * PAIR_LIST_FACTOR);
			}
		}
		const int S = d.stride;
		auto close = [&](int j, int q) {
			const double dx = kf_min_image(F[q] - F[j], L[0], iL[0]);
			const double dy = kf_min_image(F[S + q] - F[S + j], L[1], iL[1]);
			const double dz = kf_min_image(F[2 * S + q] - F[2 * S + j], L[2], iL[2]);
			return dx * dx + dy * dy + dz * dz < rl2;
		};
		const int work = on ? cnt * m2 + cnt * cnt : 0;
		int n = 0;
		for(int e = 0; e < work; e++) {
			int j, q;
			if(rec_pair(rec, s0, cnt, m0, m2, e, j, q) && close(j, q)) n++;
		}
		const int incl = wave_incl_scan(n);
		const int tot = __builtin_amdgcn_readlane(incl, 63);
		int base = 0;
		if(tot > 0) {
			if(lane == 0) base = atomicAdd(d.pl_n, tot);
			base = __builtin_amdgcn_readfirstlane(base);
		}
		if(base + tot > d.pl_cap) {
			if(lane == 0) d.pl_n[1] = 1;
			continue;
		}
		int pos = base + incl - n;
		for(int e = 0; e < work; e++) {
			int j, q;
			if(rec_pair(rec, s0, cnt, m0, m2, e, j, q) && close(j, q)) d.pl[pos++] = make_int2(j, q);
		}
	}
}

void launch_pairs_build(const DevState &d, int cus, hipStream_t s) {
	hipLaunchKernelGGL(k_pairs_reset, dim3((unsigned) ((d.R + 255) / 256)), dim3(256), 0, s, d);
	const int blocks = std::max(8, cus * 8) & ~7;
	hipLaunchKernelGGL(k_pairs_build, dim3((unsigned) blocks), dim3(256), 0, s, d);
}

// The trial energies from the close-pair list: one lane per pair, the
// separation and the bond decision as k_energy_rec forms them; a bonded pair
// adds 2 to the bond rows, an overlapping pair 1 to the overlap row.
__global__ __launch_bounds__(256) void k_energy_pairs(DevState d) {
	if(d.vslow[1] >= 0) return;  // a run-ahead sequence of attempts stopped
	const int n = min(d.pl_n[0], d.pl_cap);
	const int stride = gridDim.x * blockDim.x;
	for(int k = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; k < n; k += stride) {
		const int2 pq = d.pl[k];
		const int j = pq.x, q = pq.y;
		if((unsigned) j >= (unsigned) d.Ntot || (unsigned) q >= (unsigned) d.Ntot) continue;
		const int r = d.prep[j];
		const Rep &rp = d.reps[r];
		if(!rp.trial_active || !rp.trial_pl) continue;
		// (the early exit may read a stale flag: it only saves work)
		if(__hip_atomic_load(&d.vflag[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) continue;
		const double *F = pick_buf(d.f, rp.sel);
		const int S = d.stride;
		const double t0 = rp.trial_s[0], t1 = rp.trial_s[1], t2 = rp.trial_s[2];
		const double px = F[j] * t0, py = F[S + j] * t1, pz = F[2 * S + j] * t2;
		const double dx = kf_min_image(F[q] * t0 - px, rp.tg.box[0], rp.tg.inv_box[0]);
		const double dy = kf_min_image(F[S + q] * t1 - py, rp.tg.box[1], rp.tg.inv_box[1]);
		const double dz = kf_min_image(F[2 * S + q] * t2 - pz, rp.tg.box[2], rp.tg.inv_box[2]);
		const double d2 = dx * dx + dy * dy + dz * dz;
		const int slot = k & (NSLOT - 1);
		if(d2 < 1.) {
			__hip_atomic_store(&d.vflag[r], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			atomicAdd(ctr_at(d.ctr, r, C_OVL, slot), 1ull);
		}
		else if(d2 < rp.rc2) {
			const double c2 = rp.cosmax >= KF_FILTER_MIN_COS ? rp.cosmax * rp.cosmax : -1.;
			double O[9], P[4][3];
#pragma unroll
			for(int c = 0; c < 9; c++) O[c] = F[(3 + c) * S + j];
			kf_patches(O, P);
			if(pair_bonded(dx, dy, dz, d2, &P[0][0], c2, rp.cosmax, [&](double *Oq) {
#pragma unroll
				   for(int c = 0; c < 9; c++) Oq[c] = F[(3 + c) * S + q];
			   }))
				atomicAdd(ctr_at(d.ctr, r, C_EBONDS, slot), 2ull);
		}
	}
}

void launch_energy_pairs(const DevState &d, int cus, hipStream_t s) {
	const int blocks = std::max(8, cus * 8) & ~7;
	hipLaunchKernelGGL(k_energy_pairs, dim3((unsigned) blocks), dim3(256), 0, s, d);
}

// ------------------------------------------------------------------------
// Checkerboard invariant (a debug check; the reference is serial).  The
// particles of one colour's cells move concurrently, so no particle may lie
// within the bond range of a particle in another cell of the same colour.
// Counts the particles of replica `rep` that do, on the grid G (cells by
// cell_coord, colours by the cells' parities, separations by minimum image
// in G's box): brute force over the replica's live slots, every colour at
// once since each particle belongs to one.  One thread per slot.
__global__ void k_colour_conflicts(DevState d, int rep, int n0, int n1, int n2, unsigned long long *out) {
	const Rep &rp = d.reps[rep];
	// the grid of the current cell lists, or n0 x n1 x n2 cells over the box
	Grid G = rp.lg;
	if(n0 > 0) {
		set_grid_box(G, rp.g.box);
		G.n[0] = n0;
		G.n[1] = n1;
		G.n[2] = n2;
		G.shift[0] = G.shift[1] = G.shift[2] = 0.;
	}
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= rp.np_hw) return;
	const int j = rp.p_off + i;
	const int *PID = pick_buf(d.pid, rp.sel);
	if(PID[j] < 0) return;
	const double *F = pick_buf(d.f, rp.sel);
	const int S = d.stride;
	const double x = F[j], y = F[S + j], z = F[2 * S + j];
	auto cell = [&](double a, double b, double c, int &cx, int &cy, int &cz) {
		cx = cell_coord(a, G.shift[0], G.inv_box[0], G.n[0]);
		cy = cell_coord(b, G.shift[1], G.inv_box[1], G.n[1]);
		cz = cell_coord(c, G.shift[2], G.inv_box[2], G.n[2]);
	};
	int ix, iy, iz;
	cell(x, y, z, ix, iy, iz);
	const int col = ((ix & 1) << 2) | ((iy & 1) << 1) | (iz & 1);
	bool hit = false;
	for(int k = 0; k < rp.np_hw && !hit; k++) {
		const int q = rp.p_off + k;
		if(q == j || PID[q] < 0) continue;
		const double qx = F[q], qy = F[S + q], qz = F[2 * S + q];
		int cx, cy, cz;
		cell(qx, qy, qz, cx, cy, cz);
		if((((cx & 1) << 2) | ((cy & 1) << 1) | (cz & 1)) != col || (cx == ix && cy == iy && cz == iz)) continue;
		const double dx = kf_min_image(qx - x, G.box[0], G.inv_box[0]);
		const double dy = kf_min_image(qy - y, G.box[1], G.inv_box[1]);
		const double dz = kf_min_image(qz - z, G.box[2], G.inv_box[2]);
		hit = dx * dx + dy * dy + dz * dz < rp.rc2;
	}
	if(hit) atomicAdd(out, 1ull);
}

// nside == nullptr: the grid of the replica's current cell lists; else
// nside cells per axis over its box, no shift.  `slots`: the replica's slot
// capacity (the grid covers it; threads past its live slots leave).
void launch_colour_conflicts(const DevState &d, int rep, int slots, const int *nside, unsigned long long *out,
                             hipStream_t s) {
	const int n = std::max(slots, 1);
	hipLaunchKernelGGL(k_colour_conflicts, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, d, rep,
	                   nside ? nside[0] : 0, nside ? nside[1] : 0, nside ? nside[2] : 0, out);
}

// ------------------------------------------------------------------------
// NPT volume move (MC.c:450-513): log-volume step along one random axis

// Whether the trial of a replica can be evaluated on its current lists: the
// lists' cells, rescaled with the trial, stay at least side_min wide on
// every axis, so every pair within the bond range still lies in adjacent
// cells (the margin in side_min covers the roundings of the membership).
// Otherwise (or when asked for every trial) the trial regroups the particles
// on its own grid into the spare buffer; the host launches that rebuild only
// when some replica needs it.
__device__ __forceinline__ void choose_trial_path(const DevState &d, Rep &rp, const double tb[3]) {
	bool fast = d.npt_trial_form != 2 && rp.lists_ok;
	for(int a = 0; a < 3; a++) fast = fast && tb[a] / (double) rp.lg.n[a] >= rp.side_min;
	rp.trial_fast = fast ? 1 : 0;
	// a fast trial takes its energy from the cell records while they describe
	// the lists (the sweep's shifting rebuild wrote them)
	rp.trial_rec = (fast && rp.rec_ok && (d.npt_trial_form == 0 || d.npt_trial_form == 3)) ? 1 : 0;
	// the close-pair list while the cumulative scaling keeps unlisted pairs
	// beyond the bond range (1e-9: the roundings of the list's test)
	double smin = 1e300;
	for(int a = 0; a < 3; a++) smin = fmin(smin, rp.pl_scale[a] * rp.trial_s[a]);
	rp.trial_pl = (rp.trial_rec && d.npt_trial_form == 0 && rp.pl_ok && d.pl_n && d.pl_n[1] == 0 &&
	               smin * PAIR_LIST_FACTOR >= 1. + 1e-9) ? 1 : 0;
	if(rp.trial_pl) rp.trial_rec = 0;
	atomicOr(d.vslow, fast ? (rp.trial_pl ? (int) NEED_PAIRS : (rp.trial_rec ? (int) NEED_RECORDS : (int) NEED_PARTICLES))
	                       : (int) (NEED_REBUILD | NEED_PARTICLES));
}

__global__ void k_vol_propose(DevState d, uint64_t sweep, int k) {
	int r = blockIdx.x * blockDim.x + threadIdx.x;
	if(r >= d.R || d.vslow[1] >= 0) return;
	Rep &rp = d.reps[r];
	rp.trial_active = 0;
	rp.trial_fast = 0;
	rp.trial_rec = 0;
	rp.trial_pl = 0;
	rp.vol_apply = 0;
	d.vflag[r] = 0;
	if(rp.ensemble != 5 && !rp.lx_move) return;
	u32x4 v = rng(d.seed, r, sweep, (uint32_t) k, TAG_VOLUME, 0, 0);  // attempt k of the sweep
	if(rp.ensemble != 5) {
		// box-shape move at constant volume (MC_change_Lx, MC.c:515-531):
		// Lx += U(-1/2, 1/2) Lx_change_max, Ly = Lz = sqrt(V / Lx), rejected
		// early outside [Lyz_min, Lyz_max] (or, here, without a valid grid)
		const double Lx = rp.g.box[0] + (u01(v.x) - 0.5) * rp.lx_change_max;
		const double Lyz = Lx > 0. ? sqrt(rp.V / Lx) : -1.;
		double tb[3] = {Lx, Lyz, Lyz};
		int tn[3];
		if(!(Lyz >= rp.lyz_min && Lyz <= rp.lyz_max) ||
		   !make_grid(tb, rp.side_min, rp.min_n, rp.even, rp.cap_n, tn)) {
			atomicAdd(ctr_at(d.ctr, r, C_TRY_VOL, 0), 1ull);
			return;
		}
		for(int a = 0; a < 3; a++) rp.trial_s[a] = tb[a] / rp.g.box[a];
		set_grid_box(rp.tg, tb);
		for(int a = 0; a < 3; a++) {
			rp.tg.n[a] = tn[a];
			rp.tg.shift[a] = rp.g.shift[a] * rp.trial_s[a];
		}
		rp.trial_kind = 1;
		rp.trial_active = 1;
		choose_trial_path(d, rp, tb);
		return;
	}
	const int dir = (int) (u01(v.x) * 3.);
	const double V0 = rp.g.box[0] * rp.g.box[1] * rp.g.box[2];
	const double lnV1 = log(V0) + (u01(v.y) - 0.5) * rp.rfm;
	const double V1 = exp(lnV1);
	double side = V1;
	for(int k = 0; k < 3; k++)
		if(k != dir) side /= rp.g.box[k];
	double tb[3] = {rp.g.box[0], rp.g.box[1], rp.g.box[2]};
	const double f = side / tb[dir];
	tb[dir] = side;
	int tn[3];
	const bool ok = make_grid(tb, rp.side_min, rp.min_n, rp.even, rp.cap_n, tn);
	if(!ok) {
		// the trial box cannot hold a checkerboard grid: the move is rejected
		atomicAdd(ctr_at(d.ctr, r, C_TRY_VOL, 0), 1ull);
		return;
	}
	set_grid_box(rp.tg, tb);
	for(int k = 0; k < 3; k++) {
		rp.tg.n[k] = tn[k];
		rp.tg.shift[k] = rp.g.shift[k];
	}
	rp.trial_dir = dir;
	rp.trial_f = f;
	for(int a = 0; a < 3; a++) rp.trial_s[a] = a == dir ? f : 1.;
	rp.trial_V = V1;
	rp.trial_V0 = V0;
	rp.trial_kind = 0;
	rp.trial_active = 1;
	choose_trial_path(d, rp, tb);
}

// The trial-energy paths the box attempt in flight needs (TrialNeed bits in
// vslow, gathered by k_vol_propose), written straight into the host's pinned
// word with the attempt's sequence number above them: the host waits for
// exactly this attempt's value without a copy or a stream synchronisation.
__global__ void k_vol_publish(DevState d, int *host_word, unsigned seq) {
	if(threadIdx.x == 0) {
		const int need = __hip_atomic_load(d.vslow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		__hip_atomic_store(host_word, (int) ((seq << 4) | (unsigned) need), __ATOMIC_RELAXED,
		                   __HIP_MEMORY_SCOPE_SYSTEM);
	}
}

void launch_vol_publish(const DevState &d, int *host_word, unsigned seq, hipStream_t s) {
	hipLaunchKernelGGL(k_vol_publish, dim3(1), dim3(64), 0, s, d, host_word, seq);
}

// Run-ahead box attempts (enqueued without the host reading each attempt's
// paths): attempt k proceeds on the close-pair list alone, which serves
// every fast trial in the common case; if some replica's trial needs more,
// the sequence stops here (vslow[1] = k), and the kernels of this and the
// later attempts leave at once: the attempt's proposal stays in place and
// its paths stay in vslow[0] for the host to take over from.
__global__ void k_vol_gate(DevState d, int k) {
	if(threadIdx.x == 0 && d.vslow[1] < 0 && (d.vslow[0] & ~(int) NEED_PAIRS)) d.vslow[1] = k;
}

void launch_vol_gate(const DevState &d, int k, hipStream_t s) {
	hipLaunchKernelGGL(k_vol_gate, dim3(1), dim3(64), 0, s, d, k);
}

// The end of a run-ahead sequence: where it stopped (attempt + 1, 0 if it
// ran through) with the host's sequence number into the pinned word; the
// stop is cleared for the attempts that follow.
__global__ void k_vol_report(DevState d, int *host_word, unsigned seq) {
	if(threadIdx.x == 0) {
		const int stop = d.vslow[1];
		d.vslow[1] = -1;
		__hip_atomic_store(host_word, (int) ((seq << 8) | (unsigned) (stop + 1)), __ATOMIC_RELAXED,
		                   __HIP_MEMORY_SCOPE_SYSTEM);
	}
}

void launch_vol_report(const DevState &d, int *host_word, unsigned seq, hipStream_t s) {
	hipLaunchKernelGGL(k_vol_report, dim3(1), dim3(64), 0, s, d, host_word, seq);
}

void launch_vol_propose(const DevState &d, uint64_t sweep, int k, hipStream_t s) {
	hipLaunchKernelGGL(k_vol_propose, dim3((d.R + 63) / 64), dim3(64), 0, s, d, sweep, k);
}

__global__ void k_vol_decide(DevState d, uint64_t sweep, int k) {
	int r = blockIdx.x * blockDim.x + threadIdx.x;
	if(r >= d.R || d.vslow[1] >= 0) return;
	Rep &rp = d.reps[r];
	if(!rp.trial_active) return;
	unsigned long long eb = 0, ov = 0, db = 0;
	for(int s = 0; s < NSLOT; s++) {
		eb += *ctr_at(d.ctr, r, C_EBONDS, s);
		ov += *ctr_at(d.ctr, r, C_OVL, s);
		db += *ctr_at(d.ctr, r, C_DBONDS, s);
	}
	const long long B1 = (long long) eb / 2;
	const long long B0 = rp.bonds_base + (long long) db;
	const double dE = (double) (B0 - B1);  // E' - E with E = -bonds
	// volume: MC.c:497-499; box shape: exp(-dE/T) at constant volume (the
	// Jacobian of the rescaling is 1), or, on request, the reference's
	// exp(+dE/T) (MC.c:562), which does not keep the Boltzmann distribution
	const double arg = rp.trial_kind == 1
	                       ? (rp.lx_ref_rule ? dE / rp.T : -dE / rp.T)
	                       : -(rp.P * (rp.trial_V - rp.trial_V0) + dE) / rp.T + (double) (rp.np + 1) * log(rp.trial_f);
	u32x4 v = rng(d.seed, r, sweep, (uint32_t) k, TAG_VOLUME, 0, 0);  // attempt k of the sweep
	atomicAdd(ctr_at(d.ctr, r, C_TRY_VOL, 0), 1ull);
	if(ov == 0 && u01(v.z) < exp(arg)) {
		if(rp.trial_fast) {
			// the positions are rescaled in place (k_vol_apply); the lists keep
			// their members, their grid is rescaled with the positions
			rp.vol_apply = 1;
			for(int a = 0; a < 3; a++) rp.pl_scale[a] *= rp.trial_s[a];
			set_grid_box(rp.lg, rp.tg.box);
			for(int a = 0; a < 3; a++) rp.lg.shift[a] *= rp.trial_s[a];
		}
		else {
			// the spare buffer holds the trial, grouped on the trial grid
			// (the cell records describe the old lists)
			rp.sel ^= 1;
			rp.lg = rp.tg;
			rp.lists_ok = 1;
			rp.rec_ok = 0;
			rp.pl_ok = 0;
			rp.np_hw = rp.np;
		}
		rp.g = rp.tg;
		if(rp.trial_kind == 0) rp.V = rp.trial_V;
		rp.bonds_base = B1 - (long long) db;
		atomicAdd(ctr_at(d.ctr, r, C_ACC_VOL, 0), 1ull);
	}
	else if(!rp.trial_fast)
		rp.lists_ok = 0;  // the lists now group the spare buffer, not the current one
	rp.trial_active = 0;
}

void launch_vol_decide(const DevState &d, uint64_t sweep, int k, hipStream_t s) {
	hipLaunchKernelGGL(k_vol_decide, dim3((d.R + 63) / 64), dim3(64), 0, s, d, sweep, k);
}

// The accepted fast trials: every position component times its trial scale,
// as the trial rebuild writes it (x * 1 is x).  Clears the trial paths for
// the next attempt.
__global__ void k_vol_apply(DevState d) {
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(d.vslow[1] >= 0) return;  // a run-ahead sequence of attempts stopped
	if(j == 0) *d.vslow = 0;
	if(j >= d.Ntot) return;
	const Rep &rp = d.reps[d.prep[j]];
	if(!rp.vol_apply) return;
	double *F = pick_buf(d.f, rp.sel);
	const int S = d.stride;
	for(int c = 0; c < 3; c++) F[(size_t) c * S + j] *= rp.trial_s[c];
}

void launch_vol_apply(const DevState &d, hipStream_t s) {
	hipLaunchKernelGGL(k_vol_apply, dim3((unsigned) std::max(1, (d.Ntot + 255) / 256)), dim3(256), 0, s, d);
}

// ------------------------------------------------------------------------
// Grand-canonical insertions and removals (MC_add_remove, MC.c:342-448),
// made cell-local for the checkerboard: once per sweep, after the colour
// passes, every cell of one colour (drawn per replica and sweep; empty
// cells too) tries one insertion or one removal, each with probability 1/2.
//   insertion: a position uniform in the cell and an orientation uniform on
//     SO(3) (random_orientation, utils.c:133-169), accepted with
//     min(1, z v_c / (n_c + 1) exp(-dE/T)), dE = -(bonds of the new particle)
//   removal: a member uniform among the cell's n_c, accepted with
//     min(1, n_c / (z v_c) exp(-dE/T)), dE = +(the member's bonds)
// (MC.c:367 and MC.c:391 with the cell's volume and population in place of
// the box's).  Cells of one colour are a cell side apart, so their moves do
// not interact: each pair of moves satisfies detailed balance for
// z^N / N! exp(-E/T) on its own.  A cell's insertion takes a slot reserved
// for that cell (np + its index among the colour's cells), so the outcome
// does not depend on thread timing; a replica with fewer free slots than
// cells of one colour skips the phase and reports ERR_GC_FULL.

// Begin a GC phase: particle indices of each grand-canonical replica become
// 0 .. np - 1 in slot order.  After the sweep's passes every cell holds its
// members in order of particle index, so slot order is the same for every
// launch shape; this keeps the indices dense for the phase's insertions,
// which take indices from np on.
__global__ void k_gc_begin(DevState d, uint64_t sweep) {
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= d.Ntot) return;
	const int r = d.prep[j];
	Rep &rp = d.reps[r];
	if(!rp.gc) return;
	const int loc = j - rp.p_off;
	if(loc < rp.np) pick_buf(d.pid, rp.sel)[j] = loc;
	if(loc == 0) {
		const u32x4 v = rng(d.seed, r, sweep, 0, TAG_GC, 0, 0);
		rp.gc_colour = (int) (v.x & 7u);
		const int ncol = (rp.g.n[0] / 2) * (rp.g.n[1] / 2) * (rp.g.n[2] / 2);
		if(rp.np + ncol > rp.cap_np) {
			atomicOr(d.err, (int) ERR_GC_FULL);
			rp.gc_base = -1;
		}
		else {
			rp.gc_base = rp.np;
			rp.np_hw = rp.np + ncol;
			rp.lists_ok = 0;  // the phase adds and removes particles the lists do not show
			rp.rec_ok = 0;
		}
	}
}

// Orientation uniform on SO(3) as random_orientation builds it (utils.c:
// 133-169): three uniform unit vectors x, y, z, Gram-Schmidt in that order,
// rows (z, x, y), or (x, z, y) when that has determinant -1.
__device__ inline void gc_orientation(uint64_t seed, int r, uint64_t sweep, uint32_t cell, uint32_t col, double O[9]) {
	double v[9];
	for(int k = 0; k < 3; k++) sphere_axis(seed, r, sweep, cell, col, 0x8000u + (uint32_t) k, v + 3 * k);
	gram_schmidt_rows(v);
	const double *x = v, *y = v + 3, *z = v + 6;
	const double det = z[0] * (x[1] * y[2] - x[2] * y[1]) - z[1] * (x[0] * y[2] - x[2] * y[0]) +
	                   z[2] * (x[0] * y[1] - x[1] * y[0]);
	const double *r0 = det < 0. ? x : z, *r1 = det < 0. ? z : x;
	for(int k = 0; k < 3; k++) {
		O[k] = r0[k];
		O[3 + k] = r1[k];
		O[6 + k] = y[k];
	}
}

// Bonds of a particle at p with patches P against every particle of the 27
// cells around cell (ix, iy, iz) but slot `skip` (MC_energy, MC.c:272-309):
// T lanes, the cells shared out; every lane returns the team's count and
// whether any overlap was met.
template <int T>
__device__ inline int gc_bonds(const DevState &d, const Rep &rp, const double *F, int ix, int iy, int iz,
                               const double p[3], const double P[4][3], int skip, int lane, bool &any) {
	const Grid &G = rp.g;
	const int S = d.stride, n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
	int nb_ = 0;
	bool ov = false;
	for(int nb = lane; nb < 27; nb += T) {
		const int cx = wrapi(ix + nb / 9 - 1, n0), cy = wrapi(iy + (nb / 3) % 3 - 1, n1), cz = wrapi(iz + nb % 3 - 1, n2);
		const int cc = rp.cell_off + (cx * n1 + cy) * n2 + cz;
		for(int q = d.start[cc]; q < d.start[cc + 1]; q++) {
			if(q == skip) continue;
			double dx = kf_min_image(F[q] - p[0], G.box[0], G.inv_box[0]);
			double dy = kf_min_image(F[S + q] - p[1], G.box[1], G.inv_box[1]);
			double dz = kf_min_image(F[2 * S + q] - p[2], G.box[2], G.inv_box[2]);
			const double d2 = dx * dx + dy * dy + dz * dz;
			if(d2 < 1.) ov = true;
			else if(d2 < rp.rc2) {
				kf_unit(dx, dy, dz, d2);
				if(kf_any_p(dx, dy, dz, &P[0][0], rp.cosmax)) {
					double Oq[9];
					for(int c = 0; c < 9; c++) Oq[c] = F[(3 + c) * S + q];
					nb_ += kf_any_q(dx, dy, dz, Oq, rp.cosmax);
				}
			}
		}
	}
	any = team_any<T>(ov);
	return team_sum<T>(nb_);
}

// One team of 32 lanes per cell of the phase's colour, over every replica's
// cells (entry gi of the replicas' per-colour capacities).
__global__ __launch_bounds__(256) void k_gc_pass(DevState d, uint64_t sweep) {
	constexpr int T = 32;
	const int lane = threadIdx.x & (T - 1);
	const int gi = blockIdx.x * (blockDim.x / T) + threadIdx.x / T;
	if(gi >= d.total_act) return;
	int lo = 0, hi = d.R;
	while(hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if(d.act_off[mid] <= gi) lo = mid;
		else hi = mid;
	}
	const int r = lo;
	const Rep &rp = d.reps[r];
	if(!rp.gc || rp.gc_base < 0) return;
	const int n0 = rp.g.n[0], n1 = rp.g.n[1], n2 = rp.g.n[2];
	const int h0 = n0 / 2, h1 = n1 / 2, h2 = n2 / 2;
	const int k = gi - d.act_off[r];
	if(k >= h0 * h1 * h2) return;
	const int col = rp.gc_colour;
	const int ix = 2 * (k / (h1 * h2)) + ((col >> 2) & 1), iy = 2 * ((k / h2) % h1) + ((col >> 1) & 1),
	          iz = 2 * (k % h2) + (col & 1);
	const int cloc = (ix * n1 + iy) * n2 + iz;
	const int c = rp.cell_off + cloc;
	const int s0 = d.start[c], cnt = d.start[c + 1] - s0;
	double *F = pick_buf(d.f, rp.sel);
	int *PID = pick_buf(d.pid, rp.sel);
	const int S = d.stride;
	const u32x4 u = rng(d.seed, r, sweep, (uint32_t) cloc, TAG_GC, (uint32_t) col, 1);
	const u32x4 w = rng(d.seed, r, sweep, (uint32_t) cloc, TAG_GC, (uint32_t) col, 2);
	const double vc = (rp.g.box[0] / n0) * (rp.g.box[1] / n1) * (rp.g.box[2] / n2);
	const int cslot = gi & (NSLOT - 1);
	if(u01(u.x) < 0.5) {
		// insertion at a uniform point of the cell: the point where the
		// cell's own coordinate arithmetic puts it (a draw that rounds into a
		// neighbouring cell is rejected)
		double p[3];
		const double fr[3] = {u01(u.y), u01(u.z), u01(u.w)};
		const int ii[3] = {ix, iy, iz}, nn[3] = {n0, n1, n2};
		bool inside = true;
		for(int a = 0; a < 3; a++) {
			p[a] = rp.g.shift[a] + ((double) ii[a] + fr[a]) / (double) nn[a] * rp.g.box[a];
			inside = inside && cell_coord(p[a], rp.g.shift[a], rp.g.inv_box[a], nn[a]) == ii[a];
		}
		double O[9], P[4][3];
		gc_orientation(d.seed, r, sweep, (uint32_t) cloc, (uint32_t) col, O);
		kf_patches(O, P);
		bool ov = false;
		const int b = inside ? gc_bonds<T>(d, rp, F, ix, iy, iz, p, P, -1, lane, ov) : 0;
		if(lane != 0) return;
		atomicAdd(ctr_at(d.ctr, r, C_TRY_ADD, cslot), 1ull);
		// MC.c:367: exp(-dE/T) z V / (N + 1), dE = -b
		const double acc = exp((double) b / rp.T) * rp.activity * vc / (double) (cnt + 1);
		if(!inside || ov || !(u01(w.x) < acc)) return;
		const int slot = rp.p_off + rp.gc_base + k;
		for(int a = 0; a < 3; a++) F[(size_t) a * S + slot] = p[a];
		for(int a = 0; a < 9; a++) F[(size_t) (3 + a) * S + slot] = O[a];
		PID[slot] = rp.gc_base + k;
		atomicAdd(&d.reps[r].np, 1);
		atomicAdd(ctr_at(d.ctr, r, C_ACC_ADD, cslot), 1ull);
		if(b) atomicAdd(ctr_at(d.ctr, r, C_DBONDS, cslot), (unsigned long long) (long long) b);
	}
	else {
		if(cnt == 0) {
			if(lane == 0) atomicAdd(ctr_at(d.ctr, r, C_TRY_REM, cslot), 1ull);
			return;
		}
		// members sit in order of particle index: rank m is slot s0 + m
		const int q = s0 + (int) __umulhi(u.y, (uint32_t) cnt);
		const double p[3] = {F[q], F[S + q], F[2 * S + q]};
		double O[9], P[4][3];
		for(int a = 0; a < 9; a++) O[a] = F[(size_t) (3 + a) * S + q];
		kf_patches(O, P);
		bool ov = false;
		const int b = gc_bonds<T>(d, rp, F, ix, iy, iz, p, P, q, lane, ov);
		if(lane != 0) return;
		atomicAdd(ctr_at(d.ctr, r, C_TRY_REM, cslot), 1ull);
		// MC.c:391: exp(-dE/T) N / (V z), dE = +b
		const double acc = exp(-(double) b / rp.T) * (double) cnt / (vc * rp.activity);
		if(!(u01(w.x) < acc)) return;
		PID[q] = -1;
		atomicSub(&d.reps[r].np, 1);
		atomicAdd(ctr_at(d.ctr, r, C_ACC_REM, cslot), 1ull);
		if(b) atomicAdd(ctr_at(d.ctr, r, C_DBONDS, cslot), (unsigned long long) (long long) -b);
	}
}

void launch_gc_phase(const DevState &d, uint64_t sweep, hipStream_t s) {
	if(d.Ntot <= 0 || d.total_act <= 0) return;
	hipLaunchKernelGGL(k_gc_begin, dim3((d.Ntot + 255) / 256), dim3(256), 0, s, d, sweep);
	hipLaunchKernelGGL(k_gc_pass, dim3((d.total_act + 7) / 8), dim3(256), 0, s, d, sweep);
}

// ------------------------------------------------------------------------
// Gram-Schmidt of every orientation (MC_check_energy, MC.c:775)

__global__ void k_gram_schmidt(DevState d) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= d.Ntot) return;
	const Rep &rp = d.reps[d.prep[j]];
	if(j - rp.p_off >= rp.np) return;
	double *F = pick_buf(d.f, rp.sel);
	const int S = d.stride;
	double O[9];
#pragma unroll
	for(int c = 0; c < 9; c++) O[c] = F[(3 + c) * S + j];
	gram_schmidt_rows(O);
#pragma unroll
	for(int c = 0; c < 9; c++) F[(3 + c) * S + j] = O[c];
}

void launch_gram_schmidt(const DevState &d, hipStream_t s) {
	hipLaunchKernelGGL(k_gram_schmidt, dim3((d.Ntot + 255) / 256), dim3(256), 0, s, d);
}

}  // namespace ppd
