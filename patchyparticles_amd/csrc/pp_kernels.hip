// pp_kernels.hip -- gfx950 kernels of the Kern-Frenkel rototranslation hot path.
//
// Data layout in HBM (per context, all replicas concatenated):
//   f[2][12][stride]  double  x, y, z, O00..O22 of slot j (SoA, two buffers)
//   pid[2][stride]    int     particle index of slot j
//   cells             int     count/loc/start over the concatenated cell
//                             regions of all replicas (one region per replica,
//                             sized for its largest grid)
// Slots are sorted by cell at every rebuild; inside a cell by particle index,
// so the order (and with it the whole trajectory) depends only on the state,
// never on thread timing.
//
// Sweep = rebuild on a randomly shifted grid + 8 colour passes in random
// order.  In a pass every cell of one colour (i,j,k parities) is active; its
// neighbourhood holds no other active cell, so one team of T lanes per active
// cell can move that cell's particles while all others stay put.  A trial that
// leaves the cell is rejected, which keeps cell populations fixed during the
// pass: the team makes exactly as many sequential trials as the cell holds,
// so a sweep is exactly N attempts, as a reference sweep (MC.c:19-29).
#pragma clang fp contract(off)

#include "pp_internal.h"

namespace ppd {

// ------------------------------------------------------------------------
// helpers

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of logical blocks so that neighbouring cells share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
	int per = nb >> 3;
	return (b & 7) * per + (b >> 3);
}

template <int T>
__device__ __forceinline__ int team_sum(int v) {
#pragma unroll
	for(int o = T / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, T);
	return v;
}

template <int T>
__device__ __forceinline__ bool team_any(bool p) {
	unsigned long long m = __ballot(p);
	int off = (threadIdx.x & 63) & ~(T - 1);
	unsigned long long tm = (T == 64) ? ~0ull : ((1ull << T) - 1ull);
	return ((m >> off) & tm) != 0ull;
}

// When a team is a whole wavefront (T = 64) every team-uniform value is
// wave-uniform: readfirstlane moves it to scalar registers.
template <int T>
__device__ __forceinline__ int uni(int v) {
	if constexpr(T == 64) return __builtin_amdgcn_readfirstlane(v);
	else return v;
}

template <int T>
__device__ __forceinline__ double unid(double v) {
	if constexpr(T == 64) {
		long long b = __double_as_longlong(v);
		int lo = __builtin_amdgcn_readfirstlane((int) (b & 0xffffffffll));
		int hi = __builtin_amdgcn_readfirstlane((int) (b >> 32));
		return __longlong_as_double(((long long) hi << 32) | (unsigned) lo);
	}
	else return v;
}

template <int T>
__device__ __forceinline__ u32x4 uni4(u32x4 v) {
	return u32x4{(uint32_t) uni<T>((int) v.x), (uint32_t) uni<T>((int) v.y), (uint32_t) uni<T>((int) v.z),
	             (uint32_t) uni<T>((int) v.w)};
}

// Lanes of one team live in one wavefront: order their LDS traffic.
__device__ __forceinline__ void team_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned long long *ctr_at(unsigned long long *ctr, int r, int row, int slot) {
	return ctr + ((long long) (r * NCTR + row) * NSLOT + slot);
}

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

// Zero the energy-kernel rows (C_EBONDS, C_OVL) of every replica.
__global__ void k_zero_energy_rows(unsigned long long *ctr, int R) {
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= R * 2 * NSLOT) return;
	int r = i / (2 * NSLOT), k = i % (2 * NSLOT);
	int row = k < NSLOT ? C_EBONDS : C_OVL;
	*ctr_at(ctr, r, row, k % NSLOT) = 0ull;
}

// ------------------------------------------------------------------------
// per-sweep replica set-up: random grid shift in [0, side) per axis

__global__ void k_prep(DevState d, uint64_t sweep, long long trace_pos) {
	int r = blockIdx.x * blockDim.x + threadIdx.x;
	if(r >= d.R) return;
	Rep &rp = d.reps[r];
	// energy trace: bonded pairs as they stand after the previous sweep
	if(d.trace && trace_pos >= 0 && trace_pos < d.trace_cap) {
		unsigned long long db = 0;
		for(int k = 0; k < NSLOT; k++) db += *ctr_at(d.ctr, r, C_DBONDS, k);
		d.trace[(long long) r * d.trace_cap + trace_pos] = rp.bonds_base + (long long) db;
	}
	u32x4 v = rng(d.seed, r, sweep, 0, TAG_SHIFT, 0, 0);
	uint32_t w[3] = {v.x, v.y, v.z};
	for(int k = 0; k < 3; k++) rp.g.shift[k] = u01(w[k]) * (rp.g.box[k] / (double) rp.g.n[k]);
	rp.trial_active = 0;
}

void launch_prep(const DevState &d, uint64_t sweep, long long trace_pos, hipStream_t s) {
	hipLaunchKernelGGL(k_prep, dim3((d.R + 63) / 64), dim3(64), 0, s, d, sweep, trace_pos);
}

__global__ void k_flip(Rep *reps, int R) {
	int r = blockIdx.x * blockDim.x + threadIdx.x;
	if(r < R) reps[r].sel ^= 1;
}

void launch_flip(const DevState &d, hipStream_t s) {
	hipLaunchKernelGGL(k_flip, dim3((d.R + 63) / 64), dim3(64), 0, s, d.reps, d.R);
}

// ------------------------------------------------------------------------
// rebuild: count -> scan -> scatter -> per-cell sort -> gather

__global__ void k_count(DevState d, int trial) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= d.Ntot) return;
	const Rep &rp = d.reps[d.prep[j]];
	const bool tr = trial && rp.trial_active;
	const Grid &G = tr ? rp.tg : rp.g;
	const double *F = d.f[rp.sel];
	const int S = d.stride;
	double p[3] = {F[j], F[S + j], F[2 * S + j]};
	if(tr) p[rp.trial_dir] *= rp.trial_f;  // MC.c:475-478
	int ix = cell_coord(p[0], G.shift[0], G.inv_box[0], G.n[0]);
	int iy = cell_coord(p[1], G.shift[1], G.inv_box[1], G.n[1]);
	int iz = cell_coord(p[2], G.shift[2], G.inv_box[2], G.n[2]);
	int c = rp.cell_off + (ix * G.n[1] + iy) * G.n[2] + iz;
	d.cellid[j] = c;
	d.slot[j] = atomicAdd(&d.count[c], 1);
}

// Tile-local exclusive scan of count (and zero count for the next rebuild).
__global__ __launch_bounds__(256) void k_scan1(DevState d) {
	__shared__ int wsum[4];
	const int t = threadIdx.x;
	const int base = blockIdx.x * SCAN_TILE + t * 8;
	int v[8];
	int tot = 0;
#pragma unroll
	for(int e = 0; e < 8; e++) {
		int c = base + e;
		v[e] = c < d.C_total ? d.count[c] : 0;
		tot += v[e];
	}
	// inclusive scan of thread totals within the wavefront
	int lane = t & 63, wid = t >> 6;
	int inc = tot;
#pragma unroll
	for(int o = 1; o < 64; o <<= 1) {
		int u = __shfl_up(inc, o, 64);
		if(lane >= o) inc += u;
	}
	if(lane == 63) wsum[wid] = inc;
	__syncthreads();
	int woff = 0;
	for(int w = 0; w < wid; w++) woff += wsum[w];
	int run = woff + inc - tot;
#pragma unroll
	for(int e = 0; e < 8; e++) {
		int c = base + e;
		if(c < d.C_total) {
			d.loc[c] = run;
			d.count[c] = 0;
		}
		run += v[e];
	}
	if(t == 255) d.tile_sum[blockIdx.x] = woff + inc;
}

// Exclusive scan of the tile sums, one workgroup of 1024.
__global__ __launch_bounds__(1024) void k_scan2(DevState d) {
	__shared__ int wsum[16];
	const int t = threadIdx.x;
	const int per = (d.ntiles + 1023) / 1024;
	const int b = t * per;
	int tot = 0;
	for(int e = 0; e < per; e++)
		if(b + e < d.ntiles) tot += d.tile_sum[b + e];
	int lane = t & 63, wid = t >> 6;
	int inc = tot;
#pragma unroll
	for(int o = 1; o < 64; o <<= 1) {
		int u = __shfl_up(inc, o, 64);
		if(lane >= o) inc += u;
	}
	if(lane == 63) wsum[wid] = inc;
	__syncthreads();
	int woff = 0;
	for(int w = 0; w < wid; w++) woff += wsum[w];
	int run = woff + inc - tot;
	for(int e = 0; e < per; e++)
		if(b + e < d.ntiles) {
			int x = d.tile_sum[b + e];
			d.tile_sum[b + e] = run;
			run += x;
		}
}

__global__ void k_fix_scatter(DevState d) {
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if(t < d.C_total) d.start[t] = d.loc[t] + d.tile_sum[t / SCAN_TILE];
	else if(t == d.C_total) d.start[t] = d.Ntot;
	if(t < d.Ntot) {
		int c = d.cellid[t];
		d.idx[d.loc[c] + d.tile_sum[c / SCAN_TILE] + d.slot[t]] = t;
	}
}

// Order each cell's slots by particle index (insertion sort; cells are small).
__global__ void k_sort(DevState d) {
	int c = blockIdx.x * blockDim.x + threadIdx.x;
	if(c >= d.C_total) return;
	int s = d.start[c], e = d.start[c + 1];
	if(e - s < 2) return;
	const int *pid = d.pid[d.reps[d.prep[d.idx[s]]].sel];
	for(int a = s + 1; a < e; a++) {
		int v = d.idx[a];
		int key = pid[v];
		int b = a - 1;
		while(b >= s && pid[d.idx[b]] > key) {
			d.idx[b + 1] = d.idx[b];
			b--;
		}
		d.idx[b + 1] = v;
	}
}

__global__ void k_gather(DevState d, int trial) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= d.Ntot) return;
	const Rep &rp = d.reps[d.prep[j]];
	const bool tr = trial && rp.trial_active;
	if(trial && !tr) return;
	const int src = d.idx[j];
	const double *F = d.f[rp.sel];
	double *G = d.f[rp.sel ^ 1];
	const int S = d.stride;
#pragma unroll
	for(int c = 0; c < NCOMP; c++) {
		double v = F[c * S + src];
		if(tr && c == rp.trial_dir) v *= rp.trial_f;
		G[c * S + j] = v;
	}
	d.pid[rp.sel ^ 1][j] = d.pid[rp.sel][src];
}

void launch_rebuild(const DevState &d, int trial, hipStream_t s) {
	const int B = 256;
	unsigned gp = (unsigned) ((d.Ntot + B - 1) / B);
	hipLaunchKernelGGL(k_count, dim3(gp), dim3(B), 0, s, d, trial);
	hipLaunchKernelGGL(k_scan1, dim3(d.ntiles), dim3(256), 0, s, d);
	hipLaunchKernelGGL(k_scan2, dim3(1), dim3(1024), 0, s, d);
	long long nfix = (d.C_total + 1 > d.Ntot) ? d.C_total + 1 : d.Ntot;
	hipLaunchKernelGGL(k_fix_scatter, dim3((unsigned) ((nfix + B - 1) / B)), dim3(B), 0, s, d);
	hipLaunchKernelGGL(k_sort, dim3((unsigned) ((d.C_total + B - 1) / B)), dim3(B), 0, s, d);
	hipLaunchKernelGGL(k_gather, dim3(gp), dim3(B), 0, s, d, trial);
}

// ------------------------------------------------------------------------
// neighbour-cell enumeration shared by the pass and energy kernels:
// lane l of a team of T takes the cells nb = l/S, l/S + T/S, ... (nb in
// 0..26, dx = nb/9-1, dy = (nb/3)%3-1, dz = nb%3-1) and, inside each, the
// members m = l%S, l%S + S, ...   S = 2 for T = 64, else 1.

template <int T>
struct NbLayout {
	static constexpr int S = (T >= 54) ? 2 : 1;
	static constexpr int STEP = T / S;
	static constexpr int PER = (27 + STEP - 1) / STEP;
};

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// ------------------------------------------------------------------------
// the colour pass

struct PassArgs {
	DevState d;
	uint64_t sweep;
	int pass;
};

// per-team LDS: the cell's own particles, then the trial's patch vectors
// (old, new) and new orientation
constexpr int TEAM_LDS = CELL_CAP * NCOMP + 36;

// The 26 cells around an active cell (the cell itself is handled apart,
// from the team's LDS copy).  Lane l takes e = l/S, l/S + T/S, ...
template <int T>
struct Nb26 {
	static constexpr int S = (T >= 52) ? 2 : 1;
	static constexpr int STEP = T / S;
	static constexpr int PER = (26 + STEP - 1) / STEP;
};

// Trial state shared by the helpers of one team.
struct Trial {
	double px, py, pz, nx, ny, nz;
	double Lx, Ly, Lz, iLx, iLy, iLz, rc2, cosmax;
};

// hard-core test of the new position against q
__device__ __forceinline__ bool overlaps_new(const Trial &t, double qx, double qy, double qz) {
	double dx = kf_min_image(qx - t.nx, t.Lx, t.iLx);
	double dy = kf_min_image(qy - t.ny, t.Ly, t.iLy);
	double dz = kf_min_image(qz - t.nz, t.Lz, t.iLz);
	return (dx * dx + dy * dy + dz * dz) < 1.;
}

// bonds of q with the old and the new position; q's orientation is read
// only when one of the two passes the p-side angular test.  OWN selects the
// team's LDS copy (compile time, so no pointer ever mixes address spaces).
template <bool OWN>
__device__ __forceinline__ void bonds_old_new(const Trial &t, const double *Ps, const Base &B, const double *src,
                                              int S, int q, double qx, double qy, double qz, int &bo, int &bn) {
	double ox = kf_min_image(qx - t.px, t.Lx, t.iLx);
	double oy = kf_min_image(qy - t.py, t.Ly, t.iLy);
	double oz = kf_min_image(qz - t.pz, t.Lz, t.iLz);
	double o2 = ox * ox + oy * oy + oz * oz;
	double wx = kf_min_image(qx - t.nx, t.Lx, t.iLx);
	double wy = kf_min_image(qy - t.ny, t.Ly, t.iLy);
	double wz = kf_min_image(qz - t.nz, t.Lz, t.iLz);
	double w2 = wx * wx + wy * wy + wz * wz;
	bool po = false, pn = false;
	if(o2 < t.rc2) {
		kf_unit(ox, oy, oz, o2);
		po = kf_any_p(ox, oy, oz, Ps, t.cosmax);
	}
	if(w2 < t.rc2) {
		kf_unit(wx, wy, wz, w2);
		pn = kf_any_p(wx, wy, wz, Ps + 12, t.cosmax);
	}
	if(po || pn) {
		double Oq[9];
#pragma unroll
		for(int c = 0; c < 9; c++) Oq[c] = OWN ? src[q * NCOMP + 3 + c] : src[(3 + c) * S + q];
		if(po && kf_any_q(ox, oy, oz, Oq, B, t.cosmax)) bo++;
		if(pn && kf_any_q(wx, wy, wz, Oq, B, t.cosmax)) bn++;
	}
}

// Uniform axis on the unit sphere by Marsaglia's rejection method (the
// method of random_vector_on_sphere, utils.c:115-131), fed by Philox blocks
// (two candidate pairs per block) until a pair falls inside the unit disc.
// Each pair is accepted with probability pi/4; after AXIS_BLOCKS blocks
// (chance 0.215^128) the axis is +z or -z by a random bit, which keeps the
// axis law symmetric under a -> -a and so the proposal symmetric.
constexpr uint32_t AXIS_BLOCKS = 64;

__device__ __forceinline__ void sphere_axis(uint64_t seed, int r, uint64_t sweep, uint32_t cell, uint32_t col,
                                            uint32_t t, double a[3]) {
	for(uint32_t blk = 0; blk < AXIS_BLOCKS; blk++) {
		u32x4 v = rng(seed, r, sweep, cell | (blk << 24), TAG_AXIS, col, t);
		const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
		for(int h = 0; h < 2; h++) {
			double u1 = 1. - 2. * u01(w[2 * h]), u2 = 1. - 2. * u01(w[2 * h + 1]);
			double q = u1 * u1 + u2 * u2;
			if(q < 1.) {
				double f = 2. * __builtin_sqrt(1. - q);
				a[0] = u1 * f;
				a[1] = u2 * f;
				a[2] = 1. - 2. * q;
				return;
			}
		}
	}
	u32x4 v = rng(seed, r, sweep, cell, TAG_AXIS_FALLBACK, col, t);
	a[0] = 0.;
	a[1] = 0.;
	a[2] = (v.x & 1u) ? 1. : -1.;
}

template <int T>
__global__ __launch_bounds__(256) void k_pass(PassArgs a) {
	extern __shared__ double lds[];
	using NB = Nb26<T>;
	const DevState &d = a.d;
	const int lane = threadIdx.x & (T - 1);
	const int team = threadIdx.x / T;
	const int tpb = blockDim.x / T;
	double *own = lds + team * TEAM_LDS;
	double *Ps = own + CELL_CAP * NCOMP;  // Po[12], Pn[12], On[9]
	const int g = uni<T>(xcd_remap(blockIdx.x, gridDim.x) * tpb + team);
	if(g >= d.total_act) return;

	int lo = 0, hi = d.R;
	while(hi - lo > 1) {
		int mid = (lo + hi) >> 1;
		if(d.act_off[mid] <= g) lo = mid;
		else hi = mid;
	}
	const int r = uni<T>(lo);
	const Rep &rp = d.reps[r];
	const int n0 = uni<T>(rp.g.n[0]), n1 = uni<T>(rp.g.n[1]), n2 = uni<T>(rp.g.n[2]);
	const int h1 = n1 >> 1, h2 = n2 >> 1;
	const int local = g - d.act_off[r];
	if(local >= (n0 >> 1) * h1 * h2) return;
	const int col = uni<T>(colour_of_pass(d.seed, r, a.sweep, a.pass));
	const int ix = uni<T>(2 * (local / (h2 * h1)) + ((col >> 2) & 1));
	const int iy = uni<T>(2 * ((local / h2) % h1) + ((col >> 1) & 1));
	const int iz = uni<T>(2 * (local % h2) + (col & 1));
	const int cloc = uni<T>((ix * n1 + iy) * n2 + iz);
	const int coff = uni<T>(rp.cell_off);
	const int s0 = uni<T>(d.start[coff + cloc]);
	const int cnt = uni<T>(d.start[coff + cloc + 1]) - s0;
	if(cnt == 0) return;
	if(cnt > CELL_CAP) {
		if(lane == 0) atomicOr(d.err, (int) ERR_CELL_CAP);
		return;
	}
	double *F = d.f[uni<T>(rp.sel)];
	const int S = d.stride;

	// stage the cell's own particles in LDS; the team owns them for the pass
	for(int e = lane; e < cnt * NCOMP; e += T) {
		int m = e / NCOMP, c = e - m * NCOMP;
		own[e] = F[c * S + s0 + m];
	}
	// this lane's share of the 26 surrounding cells (fixed during the pass)
	int nb_s[NB::PER], nb_n[NB::PER];
#pragma unroll
	for(int u = 0; u < NB::PER; u++) {
		int e = lane / NB::S + u * NB::STEP;
		nb_s[u] = 0;
		nb_n[u] = 0;
		if(e < 26) {
			int nb = e < 13 ? e : e + 1;
			int cx = wrapi(ix + nb / 9 - 1, n0), cy = wrapi(iy + (nb / 3) % 3 - 1, n1), cz = wrapi(iz + nb % 3 - 1, n2);
			int cc = coff + (cx * n1 + cy) * n2 + cz;
			nb_s[u] = d.start[cc];
			nb_n[u] = d.start[cc + 1] - nb_s[u];
		}
	}
	team_sync();

	Trial tr;
	tr.Lx = unid<T>(rp.g.box[0]);
	tr.Ly = unid<T>(rp.g.box[1]);
	tr.Lz = unid<T>(rp.g.box[2]);
	tr.iLx = unid<T>(rp.g.inv_box[0]);
	tr.iLy = unid<T>(rp.g.inv_box[1]);
	tr.iLz = unid<T>(rp.g.inv_box[2]);
	tr.rc2 = unid<T>(rp.rc2);
	tr.cosmax = unid<T>(rp.cosmax);
	const double disp_max = unid<T>(rp.disp_max), theta_max = unid<T>(rp.theta_max), temp = unid<T>(rp.T);
	const double sh0 = unid<T>(rp.g.shift[0]), sh1 = unid<T>(rp.g.shift[1]), sh2 = unid<T>(rp.g.shift[2]);
	int n_try = 0, n_acc = 0, d_bonds = 0;

	for(int t = 0; t < cnt; t++) {
		u32x4 ra = uni4<T>(rng(d.seed, r, a.sweep, (uint32_t) cloc, TAG_TRIAL_A, (uint32_t) col, (uint32_t) t));
		const int k = uni<T>((int) __umulhi(ra.x, (uint32_t) cnt));
		const double *me = own + k * NCOMP;
		tr.px = unid<T>(me[0]);
		tr.py = unid<T>(me[1]);
		tr.pz = unid<T>(me[2]);
		n_try++;
		// displacement uniform in +-disp_max/2 (MC.c:317-319)
		tr.nx = unid<T>(tr.px + (u01(ra.y) - 0.5) * disp_max);
		tr.ny = unid<T>(tr.py + (u01(ra.z) - 0.5) * disp_max);
		tr.nz = unid<T>(tr.pz + (u01(ra.w) - 0.5) * disp_max);
		// a trial that leaves the cell is rejected
		if(uni<T>(cell_coord(tr.nx, sh0, tr.iLx, n0)) != ix || uni<T>(cell_coord(tr.ny, sh1, tr.iLy, n1)) != iy ||
		   uni<T>(cell_coord(tr.nz, sh2, tr.iLz, n2)) != iz)
			continue;

		// phase 1: hard cores at the new position, wavefront-ballot early exit
		bool ov = false;
		for(int m = lane; m < cnt && !ov; m += T)
			if(m != k) ov = overlaps_new(tr, own[m * NCOMP], own[m * NCOMP + 1], own[m * NCOMP + 2]);
#pragma unroll
		for(int u = 0; u < NB::PER; u++)
			for(int m = lane % NB::S; m < nb_n[u] && !ov; m += NB::S) {
				const int q = nb_s[u] + m;
				ov = overlaps_new(tr, F[q], F[S + q], F[2 * S + q]);
			}
		if(team_any<T>(ov)) continue;

		// rotation: angle uniform in [0, theta_max), axis uniform on the sphere,
		// applied in the body frame O' = O R (MC.c:321-325); lane 0 stages the
		// old and new patch vectors and O' in LDS for the whole team
		u32x4 rb = uni4<T>(rng(d.seed, r, a.sweep, (uint32_t) cloc, TAG_TRIAL_B, (uint32_t) col, (uint32_t) t));
		if(lane == 0) {
			double O[9], On[9], P[4][3], axis[3];
#pragma unroll
			for(int c = 0; c < 9; c++) O[c] = me[3 + c];
			const double theta = u01(rb.x) * theta_max;
			sphere_axis(d.seed, r, a.sweep, (uint32_t) cloc, (uint32_t) col, (uint32_t) t, axis);
			double st, ct;
			sincos(theta, &st, &ct);
			rotate_body(O, axis, st, ct, On);
			kf_patches(O, d.base, P);
#pragma unroll
			for(int i = 0; i < 4; i++)
#pragma unroll
				for(int c = 0; c < 3; c++) Ps[3 * i + c] = P[i][c];
			kf_patches(On, d.base, P);
#pragma unroll
			for(int i = 0; i < 4; i++)
#pragma unroll
				for(int c = 0; c < 3; c++) Ps[12 + 3 * i + c] = P[i][c];
#pragma unroll
			for(int c = 0; c < 9; c++) Ps[24 + c] = On[c];
		}
		team_sync();

		// phase 2: bonds at the old and the new position with one load of each
		// neighbour (the particle stays in its cell: same neighbourhood)
		int bo = 0, bn = 0;
		for(int m = lane; m < cnt; m += T)
			if(m != k) bonds_old_new<true>(tr, Ps, d.base, own, 0, m, own[m * NCOMP], own[m * NCOMP + 1],
			                               own[m * NCOMP + 2], bo, bn);
#pragma unroll
		for(int u = 0; u < NB::PER; u++)
			for(int m = lane % NB::S; m < nb_n[u]; m += NB::S) {
				const int q = nb_s[u] + m;
				bonds_old_new<false>(tr, Ps, d.base, F, S, q, F[q], F[S + q], F[2 * S + q], bo, bn);
			}
		bo = team_sum<T>(bo);
		bn = team_sum<T>(bn);
		const int db = bn - bo;  // E = -bonds, dE = -db
		const double dE = (double) (-db);
		const bool acc = (dE < 0.) || (u01(rb.w) < exp(-dE / temp));  // MC.c:332
		if(acc) {
			n_acc++;
			d_bonds += db;
			if(lane == 0) {
				double *w = own + k * NCOMP;
				w[0] = tr.nx;
				w[1] = tr.ny;
				w[2] = tr.nz;
#pragma unroll
				for(int c = 0; c < 9; c++) w[3 + c] = Ps[24 + c];
			}
		}
		team_sync();
	}
	if(n_acc > 0) {
		for(int e = lane; e < cnt * NCOMP; e += T) {
			int m = e / NCOMP, c = e - m * NCOMP;
			F[c * S + s0 + m] = own[e];
		}
	}
	if(lane == 0) {
		int slot = g & (NSLOT - 1);
		atomicAdd(ctr_at(d.ctr, r, C_TRY_ROTO, slot), (unsigned long long) n_try);
		if(n_acc) {
			atomicAdd(ctr_at(d.ctr, r, C_ACC_ROTO, slot), (unsigned long long) n_acc);
			atomicAdd(ctr_at(d.ctr, r, C_DBONDS, slot), (unsigned long long) (long long) d_bonds);
		}
	}
}

template <int T>
static void pass_T(const DevState &d, uint64_t sweep, int pass, int block, hipStream_t s) {
	int tpb = block / T;
	int nb = (d.total_act + tpb - 1) / tpb;
	nb = (nb + 7) & ~7;
	size_t lds = (size_t) tpb * TEAM_LDS * sizeof(double);
	PassArgs a{d, sweep, pass};
	hipLaunchKernelGGL(k_pass<T>, dim3(nb), dim3(block), lds, s, a);
}

size_t pass_lds_bytes(int T, int block) { return (size_t) (block / T) * TEAM_LDS * sizeof(double); }

void launch_pass(const DevState &d, uint64_t sweep, int pass, int T, int block, hipStream_t s) {
	switch(T) {
	case 8: pass_T<8>(d, sweep, pass, block, s); break;
	case 16: pass_T<16>(d, sweep, pass, block, s); break;
	case 64: pass_T<64>(d, sweep, pass, block, s); break;
	default: pass_T<32>(d, sweep, pass, block, s); break;
	}
}

// ------------------------------------------------------------------------
// energy: one team per slot, MC_energy semantics (MC.c:272-309): bonds of
// the particle with every neighbour, 0 and an overlap flag on any overlap.

template <int T>
__global__ __launch_bounds__(256) void k_energy(DevState d, int trial, int per_particle) {
	using NB = NbLayout<T>;
	const int lane = threadIdx.x & (T - 1);
	const int team = threadIdx.x / T;
	const int j = blockIdx.x * (blockDim.x / T) + team;
	if(j >= d.Ntot) return;
	const int r = d.prep[j];
	const Rep &rp = d.reps[r];
	const bool tr = trial && rp.trial_active;
	if(trial && !tr) return;
	const Grid &G = tr ? rp.tg : rp.g;
	const int b = tr ? (rp.sel ^ 1) : rp.sel;
	const double *F = d.f[b];
	const int S = d.stride;
	const double px = F[j], py = F[S + j], pz = F[2 * S + j];
	double O[9];
#pragma unroll
	for(int c = 0; c < 9; c++) O[c] = F[(3 + c) * S + j];
	double P[4][3];
	kf_patches(O, d.base, P);
	const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
	const int ix = cell_coord(px, G.shift[0], G.inv_box[0], n0);
	const int iy = cell_coord(py, G.shift[1], G.inv_box[1], n1);
	const int iz = cell_coord(pz, G.shift[2], G.inv_box[2], n2);
	const double Lx = G.box[0], Ly = G.box[1], Lz = G.box[2];
	const double iLx = G.inv_box[0], iLy = G.inv_box[1], iLz = G.inv_box[2];
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
			double dx = kf_min_image(F[q] - px, Lx, iLx);
			double dy = kf_min_image(F[S + q] - py, Ly, iLy);
			double dz = kf_min_image(F[2 * S + q] - pz, Lz, iLz);
			double d2 = dx * dx + dy * dy + dz * dz;
			if(d2 < 1.) ov = true;
			else if(d2 < rp.rc2) {
				kf_unit(dx, dy, dz, d2);
				if(kf_any_p(dx, dy, dz, &P[0][0], rp.cosmax)) {
					double Oq[9];
#pragma unroll
					for(int c = 0; c < 9; c++) Oq[c] = F[(3 + c) * S + q];
					if(kf_any_q(dx, dy, dz, Oq, d.base, rp.cosmax)) nbond++;
				}
			}
		}
	}
	nbond = team_sum<T>(nbond);
	const bool any = team_any<T>(ov);
	if(lane == 0) {
		int res = any ? 0 : nbond;
		if(per_particle) {
			int i = rp.p_off + d.pid[b][j];
			d.out_bonds[i] = res;
			d.out_ovl[i] = any ? 1 : 0;
		}
		int slot = j & (NSLOT - 1);
		if(res) atomicAdd(ctr_at(d.ctr, r, C_EBONDS, slot), (unsigned long long) res);
		if(any) atomicAdd(ctr_at(d.ctr, r, C_OVL, slot), 1ull);
	}
}

void launch_energy(const DevState &d, int trial, int per_particle, int T, hipStream_t s) {
	hipLaunchKernelGGL(k_zero_energy_rows, dim3((d.R * 2 * NSLOT + 255) / 256), dim3(256), 0, s, d.ctr, d.R);
	const int block = 256;
	switch(T) {
	case 8:
		hipLaunchKernelGGL(k_energy<8>, dim3((d.Ntot + block / 8 - 1) / (block / 8)), dim3(block), 0, s, d, trial,
		                   per_particle);
		break;
	case 16:
		hipLaunchKernelGGL(k_energy<16>, dim3((d.Ntot + block / 16 - 1) / (block / 16)), dim3(block), 0, s, d,
		                   trial, per_particle);
		break;
	case 64:
		hipLaunchKernelGGL(k_energy<64>, dim3((d.Ntot + block / 64 - 1) / (block / 64)), dim3(block), 0, s, d,
		                   trial, per_particle);
		break;
	default:
		hipLaunchKernelGGL(k_energy<32>, dim3((d.Ntot + block / 32 - 1) / (block / 32)), dim3(block), 0, s, d,
		                   trial, per_particle);
		break;
	}
}

// ------------------------------------------------------------------------
// NPT volume move (MC.c:450-513): log-volume step along one random axis

__global__ void k_vol_propose(DevState d, uint64_t sweep) {
	int r = blockIdx.x * blockDim.x + threadIdx.x;
	if(r >= d.R) return;
	Rep &rp = d.reps[r];
	rp.trial_active = 0;
	if(rp.ensemble != 5) return;
	u32x4 v = rng(d.seed, r, sweep, 0, TAG_VOLUME, 0, 0);
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
	rp.trial_V = V1;
	rp.trial_V0 = V0;
	rp.trial_active = 1;
}

void launch_vol_propose(const DevState &d, uint64_t sweep, hipStream_t s) {
	hipLaunchKernelGGL(k_vol_propose, dim3((d.R + 63) / 64), dim3(64), 0, s, d, sweep);
}

__global__ void k_vol_decide(DevState d, uint64_t sweep) {
	int r = blockIdx.x * blockDim.x + threadIdx.x;
	if(r >= d.R) return;
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
	const double dV = rp.trial_V - rp.trial_V0;
	const double arg = -(rp.P * dV + dE) / rp.T + (double) (rp.np + 1) * log(rp.trial_f);
	u32x4 v = rng(d.seed, r, sweep, 0, TAG_VOLUME, 0, 0);
	atomicAdd(ctr_at(d.ctr, r, C_TRY_VOL, 0), 1ull);
	if(ov == 0 && u01(v.z) < exp(arg)) {
		rp.sel ^= 1;
		rp.g = rp.tg;
		rp.V = rp.trial_V;
		rp.bonds_base = B1 - (long long) db;
		atomicAdd(ctr_at(d.ctr, r, C_ACC_VOL, 0), 1ull);
	}
	rp.trial_active = 0;
}

void launch_vol_decide(const DevState &d, uint64_t sweep, hipStream_t s) {
	hipLaunchKernelGGL(k_vol_decide, dim3((d.R + 63) / 64), dim3(64), 0, s, d, sweep);
}

// ------------------------------------------------------------------------
// Gram-Schmidt of every orientation (MC_check_energy, MC.c:775)

__global__ void k_gram_schmidt(DevState d) {
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= d.Ntot) return;
	double *F = d.f[d.reps[d.prep[j]].sel];
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
