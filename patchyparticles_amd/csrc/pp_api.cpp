// pp_api.cpp -- the C ABI of libpp_hip.so (declared in include/pp.h).
//
// A context owns one HIP stream and the device state of one or more
// replicas; every call is synchronous for the caller.  The device copy of the
// replica table is authoritative (kernels update buffer selectors, grids,
// volumes and tracked energies); host code reads it back before changing it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pp.h"
#include "pp_internal.h"
#include "pp_net.h"

using namespace ppd;

namespace {

thread_local std::string g_err;

pp_status fail(pp_status s, const char *fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	g_err = buf;
	return s;
}

#define HIPCHECK(call)                                                                                    \
	do {                                                                                                  \
		hipError_t e_ = (call);                                                                           \
		if(e_ != hipSuccess) return fail(PP_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));      \
	} while(0)

// normalize() of the tetrahedral base vectors (system.c:20-31, utils.c:31-36)
Base host_base_patches() {
	Base B;
	const double h = 0.5 / std::sqrt(3.);
	const int sgn[4][3] = {{-1, -1, 1}, {1, -1, -1}, {1, 1, 1}, {-1, 1, -1}};
	for(int i = 0; i < 4; i++) {
		double v[3];
		for(int k = 0; k < 3; k++) v[k] = sgn[i][k] > 0 ? h : -h;
		double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
		for(int k = 0; k < 3; k++) B.b[i][k] = v[k] / len;
	}
	return B;
}

}  // namespace

// the error message of the calling thread, set by the other translation units
pp_status pp_set_error(pp_status s, const char *msg) {
	g_err = msg;
	return s;
}

struct pp_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	int R = 0;
	std::vector<Rep> reps;  // host copy (refreshed from the device when needed)
	std::vector<int> np, p_off, act_off;
	DevState d{};
	int T = 32, block = 256;
	int coop_blocks = 0;  // > 0: the passes of a sweep run as one cooperative launch of that many workgroups
	bool dd = false;      // the passes of a sweep run as one dependency-driven launch
	bool colour_check = false;  // live checkerboard checks in every sweep
	long long colour_checks = 0;  // replica checks made since the last reset
	unsigned dd_mark = 0;  // the mark the cells of the last dependency-driven sweep published
	long long sweep = 0;
	bool any_npt = false;
	bool any_gc = false;  // a replica makes grand-canonical insertions and removals
	bool any_lx = false;  // a replica makes box-shape moves
	long long trace_len = -1;  // entries written to d.trace; -1 = tracing off
	long long trace_from = 0;  // sweep index at which tracing began
	// timing
	bool timing = false;   // events around the launches being enqueued
	int timing_every = 0;  // sweeps between timed sweeps (0: off)
	std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;  // (kind, events)
	long long t_launch[5] = {0, 0, 0, 0, 0};
	double t_ms[5] = {0, 0, 0, 0, 0};
	// learned proposal: the attached network (not owned) and the share of
	// colour passes that use it
	pp_net *net = nullptr;
	double learned_fraction = 0.;
	std::vector<void *> allocs;
	// pinned: (sequence number << 4) | the trial-energy paths the box attempt
	// in flight needs (TrialNeed bits), written by the device (k_vol_publish)
	int *h_vslow = nullptr;       // [2] pinned: an attempt's trial paths; where a run-ahead sequence stopped
	int *d_vslow_host = nullptr;  // its device address
	bool vol_run_ahead = true;    // box attempts enqueued on the close-pair list, one host wait per step
	unsigned vol_seq = 0;
	// asynchronous snapshots (pp_snapshot_start / _finish): a device staging
	// area and pinned host copies of one replica's slots, copied on their own
	// stream
	hipStream_t copy_stream = nullptr;
	hipEvent_t snap_ready = nullptr, snap_done = nullptr;
	double *snap_dev = nullptr, *snap_host = nullptr;  // NCOMP x snap_cap
	int *snap_dev_pid = nullptr, *snap_host_pid = nullptr;
	int snap_cap = 0;
	bool snap_pending = false;
	int snap_n = 0;
	bool snap_gc = false;
	double snap_box[3] = {0., 0., 0.};
	int cus = 256;           // compute units of the device

	~pp_ctx() {
		for(auto &p : pending) {
			(void) hipEventDestroy(p.second.first);
			(void) hipEventDestroy(p.second.second);
		}
		if(copy_stream) (void) hipStreamSynchronize(copy_stream);
		for(void *p : allocs) (void) hipFree(p);
		if(h_vslow) (void) hipHostFree(h_vslow);
		if(snap_host) (void) hipHostFree(snap_host);
		if(snap_host_pid) (void) hipHostFree(snap_host_pid);
		if(snap_ready) (void) hipEventDestroy(snap_ready);
		if(snap_done) (void) hipEventDestroy(snap_done);
		if(copy_stream) (void) hipStreamDestroy(copy_stream);
		if(stream) (void) hipStreamDestroy(stream);
	}
};

namespace {

enum { TK_PASS = 0, TK_REBUILD = 1, TK_ENERGY = 2, TK_VOLUME = 3, TK_LEARN = 4, TK_KINDS = 5 };

template <typename Tp>
pp_status dev_alloc(pp_ctx *c, Tp **p, size_t n) {
	void *q = nullptr;
	hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(Tp));
	if(e != hipSuccess) return fail(PP_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(Tp), hipGetErrorString(e));
	c->allocs.push_back(q);
	*p = (Tp *) q;
	return PP_OK;
}

#define TRY(x)                          \
	do {                                \
		pp_status s_ = (x);             \
		if(s_ != PP_OK) return s_;      \
	} while(0)

pp_status pull_reps(pp_ctx *c) {
	HIPCHECK(hipMemcpyAsync(c->reps.data(), c->d.reps, sizeof(Rep) * c->R, hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	return PP_OK;
}

pp_status push_reps(pp_ctx *c) {
	HIPCHECK(hipMemcpyAsync(c->d.reps, c->reps.data(), sizeof(Rep) * c->R, hipMemcpyHostToDevice, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	return PP_OK;
}

pp_status pull_ctr(pp_ctx *c, std::vector<unsigned long long> &h) {
	h.resize((size_t) c->R * NCTR * NSLOT);
	HIPCHECK(hipMemcpyAsync(h.data(), c->d.ctr, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
	                        c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	return PP_OK;
}

long long ctr_sum(const std::vector<unsigned long long> &h, int r, int row) {
	unsigned long long s = 0;
	for(int k = 0; k < NSLOT; k++) s += h[((size_t) r * NCTR + row) * NSLOT + k];
	return (long long) s;
}

pp_status check_err(pp_ctx *c) {
	int e = 0;
	HIPCHECK(hipMemcpyAsync(&e, c->d.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	HIPCHECK(hipGetLastError());
	if(e & ERR_CELL_CAP) return fail(PP_ERR_CELL_CAP, "a sweep cell holds more than %d particles", CELL_CAP);
	if(e & ERR_BOX) return fail(PP_ERR_BOX, "box too small for the cell grid");
	if(e & ERR_SCAN) return fail(PP_ERR_HIP, "cell scan look-back did not complete");
	if(e & ERR_GC_FULL)
		return fail(PP_ERR_ARG, "The system contains too many particles for GC_N_max: a grand-canonical phase needs "
		                        "N plus the cells of one colour in free slots, increase GC_N_max");
	if(e & ERR_DEPWAIT) {
		launch_fill_i32(c->d.err, 0, 1, c->stream);
		HIPCHECK(hipStreamSynchronize(c->stream));
		return fail(PP_ERR_HIP, "a neighbour wait of the dependency-driven sweep kernel timed out");
	}
	if(e & ERR_BARRIER) {
		launch_fill_i32((int *) c->d.bar, 0, 2, c->stream);
		launch_fill_i32(c->d.err, 0, 1, c->stream);
		HIPCHECK(hipStreamSynchronize(c->stream));
		return fail(PP_ERR_HIP, "a grid barrier of the sweep kernel timed out");
	}
	return PP_OK;
}

// Timing is best effort: a launch whose events cannot be created or recorded
// is simply not timed.
void tic(pp_ctx *c, int kind, hipEvent_t *a) {
	(void) kind;
	*a = nullptr;
	if(!c->timing) return;
	if(hipEventCreate(a) != hipSuccess) {
		*a = nullptr;
		return;
	}
	if(hipEventRecord(*a, c->stream) != hipSuccess) {
		(void) hipEventDestroy(*a);
		*a = nullptr;
	}
}

void toc(pp_ctx *c, int kind, hipEvent_t a) {
	if(!c->timing || !a) return;
	hipEvent_t b = nullptr;
	if(hipEventCreate(&b) != hipSuccess || hipEventRecord(b, c->stream) != hipSuccess) {
		if(b) (void) hipEventDestroy(b);
		(void) hipEventDestroy(a);
		return;
	}
	c->pending.push_back({kind, {a, b}});
}

pp_status collect_timing(pp_ctx *c) {
	if(c->pending.empty()) return PP_OK;
	HIPCHECK(hipStreamSynchronize(c->stream));
	for(auto &p : c->pending) {
		float ms = 0.f;
		if(hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
			c->t_ms[p.first] += ms;
			c->t_launch[p.first] += 1;
		}
		(void) hipEventDestroy(p.second.first);
		(void) hipEventDestroy(p.second.second);
	}
	c->pending.clear();
	return PP_OK;
}

// rebuild every replica on its grid g (zero-cost for the trajectory: the
// in-cell order after a rebuild depends only on particle indices)
void rebuild_now(pp_ctx *c) {
	hipEvent_t a{};
	tic(c, TK_REBUILD, &a);
	launch_rebuild(c->d, REB_KEEP, 0, -1, c->stream);
	toc(c, TK_REBUILD, a);
}

pp_status energy_now(pp_ctx *c, int per_particle) {
	rebuild_now(c);
	hipEvent_t a{};
	tic(c, TK_ENERGY, &a);
	launch_energy(c->d, 0, per_particle, c->stream);
	toc(c, TK_ENERGY, a);
	HIPCHECK(hipStreamSynchronize(c->stream));
	return PP_OK;
}

// The particle indices of replica rep's particles in ascending order, and
// the slot-order copy they were read in.  After a rebuild a replica's
// particles fill its first np slots; a grand-canonical replica's indices are
// unique but, after an insertion, not 0 .. np - 1: callers see its particles
// in index order, renumbered 0 .. np - 1.
pp_status slot_indices(pp_ctx *c, int rep, std::vector<int> &in_slots, std::vector<int> &sorted) {
	const Rep &rp = c->reps[rep];
	in_slots.assign(std::max(rp.np, 1), 0);
	if(rp.np)
		HIPCHECK(hipMemcpy(in_slots.data(), c->d.pid[rp.sel] + rp.p_off, sizeof(int) * rp.np, hipMemcpyDeviceToHost));
	in_slots.resize(rp.np);
	sorted = in_slots;
	std::sort(sorted.begin(), sorted.end());
	for(int i = 0; i < rp.np; i++)
		if(sorted[i] < 0 || (i > 0 && sorted[i] == sorted[i - 1]) || (!rp.gc && sorted[i] != i))
			return fail(PP_ERR_HIP, "corrupt particle index %d", sorted[i]);
	return PP_OK;
}

int rank_of(const std::vector<int> &sorted, int id) {
	return (int) (std::lower_bound(sorted.begin(), sorted.end(), id) - sorted.begin());
}

pp_status validate_params(const pp_params &p, int r) {
	if(p.ensemble != PP_NVT && p.ensemble != PP_NPT && p.ensemble != PP_GC)
		return fail(PP_ERR_UNSUPPORTED, "replica %d: Ensemble %d not supported (0 = NVT, 1 = GC, 5 = NPT)", r,
		            p.ensemble);
	if(p.ensemble == PP_GC && !(p.activity > 0.)) return fail(PP_ERR_ARG, "replica %d: Activity must be > 0", r);
	if(!(p.temperature > 0.)) return fail(PP_ERR_ARG, "replica %d: Temperature must be > 0", r);
	if(!(p.kf_delta > 0.)) return fail(PP_ERR_ARG, "replica %d: KF_delta must be > 0", r);
	if(!(p.disp_max >= 0.) || !(p.theta_max >= 0.)) return fail(PP_ERR_ARG, "replica %d: negative move size", r);
	if(p.ensemble == PP_NPT && !(p.rescale_factor_max >= 0.))
		return fail(PP_ERR_ARG, "replica %d: rescale_factor_max must be >= 0", r);
	return PP_OK;
}

// common constructor: energy_only selects grids of >= 3 cells of any parity
pp_status create_impl(int R, const pp_params *prm, const pp_options *opts, const int *n, const double *r,
                      const double *O, const double *box, bool energy_only, pp_ctx **out) {
	if(!out) return fail(PP_ERR_ARG, "out is NULL");
	*out = nullptr;
	if(R <= 0 || !prm || !n || !box) return fail(PP_ERR_ARG, "bad replica arguments");
	pp_options o{};
	if(opts) o = *opts;
	if(o.threads_per_cell == 0) o.threads_per_cell = 32;
	if(o.block_size == 0) {
		// the largest workgroup whose teams' LDS fits
		o.block_size = 256;
		while(o.block_size > 64 && pass_lds_bytes(o.threads_per_cell, o.block_size) > MAX_LDS_BYTES)
			o.block_size /= 2;
	}
	const int Tv = o.threads_per_cell;
	if(Tv != 16 && Tv != 32 && Tv != 64) return fail(PP_ERR_ARG, "threads_per_cell must be 16, 32 or 64");
	if(o.block_size % 64 || o.block_size < 64 || o.block_size > 256)
		return fail(PP_ERR_ARG, "block_size must be 64, 128, 192 or 256");
	if(o.launch < PP_LAUNCH_DEFAULT || o.launch > PP_LAUNCH_DEPS) return fail(PP_ERR_ARG, "launch must be 0, 1, 2 or 3");
	if(o.launch == PP_LAUNCH_COOP && Tv == 16)
		return fail(PP_ERR_ARG, "the cooperative launch form takes 32 or 64 threads per cell");
	if(o.launch == PP_LAUNCH_DEPS && Tv == 16)
		return fail(PP_ERR_ARG, "the dependency-driven launch form takes 32 or 64 threads per cell");
	if(o.trials_per_cell < 0 || o.trials_per_cell > MAX_TRIALS_PER_CELL)
		return fail(PP_ERR_ARG, "trials_per_cell must be in [0, %d]", MAX_TRIALS_PER_CELL);
	if(pass_lds_bytes(Tv, o.block_size) > MAX_LDS_BYTES)
		return fail(PP_ERR_ARG, "threads_per_cell %d with block_size %d needs %zu bytes of LDS (max %zu)", Tv,
		            o.block_size, pass_lds_bytes(Tv, o.block_size), MAX_LDS_BYTES);

	// slots per replica: its particles, or for a grand-canonical replica the
	// most it may hold (GC_N_max)
	std::vector<int> cap(R);
	long long Ntot = 0;
	for(int k = 0; k < R; k++) {
		if(n[k] < 0) return fail(PP_ERR_ARG, "negative particle count");
		TRY(validate_params(prm[k], k));
		cap[k] = n[k];
		if(prm[k].ensemble == PP_GC) {
			if(prm[k].gc_n_max < n[k])
				return fail(PP_ERR_ARG, "replica %d: GC_N_max (%d) is below the initial number of particles (%d)", k,
				            prm[k].gc_n_max, n[k]);
			cap[k] = prm[k].gc_n_max;
		}
		Ntot += cap[k];
	}
	long long nsum = 0;
	for(int k = 0; k < R; k++) nsum += n[k];
	if(Ntot > (1LL << 30)) return fail(PP_ERR_ARG, "too many particles");
	if(nsum > 0 && (!r || !O)) return fail(PP_ERR_ARG, "positions/orientations are NULL");

	int ndev = 0;
	if(hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(PP_ERR_HIP, "no HIP device available");
	if(o.device < 0 || o.device >= ndev) return fail(PP_ERR_ARG, "device %d out of range", o.device);

	pp_ctx *c = new pp_ctx();
	c->device = o.device;
	c->R = R;
	c->T = Tv;
	c->block = o.block_size;
	c->reps.resize(R);
	c->np.assign(cap.begin(), cap.end());  // a grand-canonical replica: the most it may hold
	c->p_off.resize(R);
	c->act_off.resize(R + 1);
	auto bail = [&](pp_status s) {
		delete c;
		return s;
	};
	if(hipSetDevice(o.device) != hipSuccess) return bail(fail(PP_ERR_HIP, "hipSetDevice failed"));
	if(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
		return bail(fail(PP_ERR_HIP, "hipStreamCreate failed"));

	long long poff = 0, coff = 0, aoff = 0;
	for(int k = 0; k < R; k++) {
		Rep &rp = c->reps[k];
		memset(&rp, 0, sizeof(Rep));
		const pp_params &p = prm[k];
		const double *b = box + 3 * k;
		for(int a = 0; a < 3; a++)
			if(!(b[a] > 0.)) return bail(fail(PP_ERR_ARG, "replica %d: box side %d is not positive", k, a));
		const double rcut = 1. + p.kf_delta;
		rp.rc2 = rcut * rcut;  // SQR(1. + delta) (system.c:86)
		rp.cosmax = p.kf_cosmax;
		rp.T = p.temperature;
		for(int e = 0; e < ACC_TAB; e++) rp.acc_tab[e] = exp(-(double) e / p.temperature);  // MC.c:332
		rp.disp_max = p.disp_max;
		rp.theta_max = p.theta_max;
		rp.P = p.pressure;
		rp.rfm = p.rescale_factor_max;
		rp.ensemble = p.ensemble;
		rp.min_n = energy_only ? 3 : 4;
		rp.even = energy_only ? 0 : 1;
		// a small margin keeps cells strictly wider than the bond range
		// against the rounding of the cell-coordinate arithmetic
		double side = rcut * (1. + 1e-9);
		if(!energy_only && o.cell_occupancy > 0. && n[k] > 0) {
			double V = b[0] * b[1] * b[2];
			side = std::max(side, std::cbrt(o.cell_occupancy * V / n[k]));
		}
		rp.side_min = side;
		int gn[3];
		if(!make_grid(b, side, rp.min_n, rp.even, nullptr, gn))
			return bail(fail(PP_ERR_BOX,
			                 "replica %d: box (%g, %g, %g) holds fewer than %d cells of side %g per axis", k, b[0],
			                 b[1], b[2], rp.min_n, side));
		// a team stages at most CELL_CAP members: a grid whose mean occupancy
		// is above that cannot run (a lower one may still fill single cells
		// beyond it, which the passes report as PP_ERR_CELL_CAP)
		if(!energy_only && (double) n[k] > (double) CELL_CAP * gn[0] * gn[1] * gn[2])
			return bail(fail(PP_ERR_CELL_CAP, "replica %d: %d particles in %d sweep cells exceed %d per cell", k, n[k],
			                 gn[0] * gn[1] * gn[2], CELL_CAP));
		for(int a = 0; a < 3; a++) {
			int cap = gn[a];
			if(p.ensemble == PP_NPT) cap = (gn[a] + std::max(2, gn[a] / 8) + 1) & ~1;
			rp.cap_n[a] = cap;
			rp.g.n[a] = gn[a];
			rp.g.shift[a] = 0.;
		}
		set_grid_box(rp.g, b);
		rp.tg = rp.g;
		rp.V = b[0] * b[1] * b[2];
		rp.p_off = (int) poff;
		rp.np = n[k];
		rp.cap_np = cap[k];
		rp.np_hw = n[k];
		rp.gc = p.ensemble == PP_GC ? 1 : 0;
		rp.activity = p.activity;
		rp.cell_off = (int) coff;
		rp.cell_cap = rp.cap_n[0] * rp.cap_n[1] * rp.cap_n[2];
		rp.act_cap = (rp.cap_n[0] / 2) * (rp.cap_n[1] / 2) * (rp.cap_n[2] / 2);
		if(rp.cell_cap >= (1 << 24))
			return bail(fail(PP_ERR_BOX, "replica %d: %d cells exceed the 2^24 cells one replica may hold", k,
			                 rp.cell_cap));
		rp.sel = 0;
		c->p_off[k] = (int) poff;
		c->act_off[k] = (int) aoff;
		poff += cap[k];
		coff += rp.cell_cap + 1;  // and the parking cell of free slots
		aoff += (long long) (rp.cap_n[0] / 2) * (rp.cap_n[1] / 2) * (rp.cap_n[2] / 2);
		if(p.ensemble == PP_NPT) c->any_npt = true;
		if(p.ensemble == PP_GC) c->any_gc = true;
		if(coff > (1LL << 30)) return bail(fail(PP_ERR_NOMEM, "cell grid too large"));
	}
	c->act_off[R] = (int) aoff;

	DevState &d = c->d;
	d.Ntot = (int) Ntot;
	d.stride = (int) std::max<long long>(Ntot, 1);
	d.R = R;
	d.C_total = (int) coff;
	d.scan_tile = (d.C_total + SCAN_TILE_SMALL - 1) / SCAN_TILE_SMALL <= SCAN_SMALL_MAX_TILES ? SCAN_TILE_SMALL
	                                                                                   : SCAN_TILE_LARGE;
	d.ntiles = (d.C_total + d.scan_tile - 1) / d.scan_tile;
	d.total_act = (int) aoff;
	{
		// the kernels carry the base patches as compile-time constants
		const Base hb = host_base_patches();
		for(int i = 0; i < 4; i++)
			for(int k = 0; k < 3; k++)
				if(hb.b[i][k] != kf_base(i, k))
					return bail(fail(PP_ERR_HIP, "base patch (%d,%d) computed on the host differs from the kernels'", i, k));
	}
	d.seed = o.seed;
	d.trials_per_cell = o.trials_per_cell;
	if(o.launch == PP_LAUNCH_COOP) {
		// the passes of a sweep as one cooperative launch: as many workgroups
		// as the lists of one pass can use, at most as many as stay resident
		int coop = 0, cus = 0;
		if(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, o.device) != hipSuccess) coop = 0;
		if(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, o.device) != hipSuccess) cus = 0;
		const int tpb = o.block_size / Tv;
		const int need = std::max(1, (d.total_act + tpb - 1) / tpb);
		const int fit = (coop && cus > 0) ? sweep_blocks_max(Tv, o.block_size, cus) : 0;
		if(fit <= 0)
			return bail(fail(PP_ERR_UNSUPPORTED, "device %d cannot launch the sweep kernel cooperatively", o.device));
		c->coop_blocks = std::min(need, fit);
	}
	c->dd = o.launch == PP_LAUNCH_DEPS;
	pp_status st;
	if((st = dev_alloc(c, &d.f[0], (size_t) NCOMP * d.stride)) || (st = dev_alloc(c, &d.f[1], (size_t) NCOMP * d.stride)) ||
	   (st = dev_alloc(c, &d.pid[0], d.stride)) || (st = dev_alloc(c, &d.pid[1], d.stride)) ||
	   (st = dev_alloc(c, &d.reps, R)) || (st = dev_alloc(c, (int **) &d.prep, d.stride)) ||
	   (st = dev_alloc(c, (int **) &d.act_off, R + 1)) || (st = dev_alloc(c, &d.count, d.C_total + 1)) ||
	   (st = dev_alloc(c, &d.start, d.C_total + 1)) || (st = dev_alloc(c, &d.slot, d.stride)) ||
	   (st = dev_alloc(c, &d.cellid, d.stride)) || (st = dev_alloc(c, &d.scan_status, std::max(d.ntiles, 1))) ||
	   (st = dev_alloc(c, &d.tile_sum, std::max(d.ntiles, 1))) ||
	   (st = dev_alloc(c, &d.scan_ticket, 2)) ||
	   (st = dev_alloc(c, &d.clist, d.C_total + 1)) || (st = dev_alloc(c, &d.ccount, (size_t) R * 8)) ||
	   (st = dev_alloc(c, &d.pfx, (size_t) 8 * (R + 1))) ||
	   (st = dev_alloc(c, &d.rec, (size_t) 8 * std::max(d.total_act, 1) * REC_W)) ||
	   (st = dev_alloc(c, &d.ctr, (size_t) R * NCTR * NSLOT)) || (st = dev_alloc(c, &d.err, 1)) ||
	   (st = dev_alloc(c, &d.out_bonds, d.stride)) || (st = dev_alloc(c, &d.out_ovl, d.stride)) ||
	   (st = dev_alloc(c, &d.bar, 2)) || (st = dev_alloc(c, &d.vflag, R)) || (st = dev_alloc(c, &d.vslow, 2)) ||
	   (st = dev_alloc(c, &d.done, d.C_total + 1)) || (st = dev_alloc(c, &d.ddq, 1)) ||
	   (st = dev_alloc(c, &d.ccheck, 2)) || (st = dev_alloc(c, &d.ovf, std::max(d.total_act, 1))) ||
	   (st = dev_alloc(c, &d.ovf_n, 8)))
		return bail(st);
	if(c->any_npt || c->any_lx) {
		// the close-pair list of the box trials: 8 pairs per slot (a
		// hard-sphere particle has at most about 14 neighbours within 1.1
		// bond ranges; a fuller list overflows to the cell records)
		d.pl_cap = (int) std::min<long long>(8LL * d.stride, 0x7fffff00LL);
		if((st = dev_alloc(c, &d.pl, (size_t) d.pl_cap)) || (st = dev_alloc(c, &d.pl_n, 2))) return bail(st);
	}
	if(hipHostMalloc((void **) &c->h_vslow, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess) c->h_vslow = nullptr;
	if(!c->h_vslow || hipHostGetDevicePointer((void **) &c->d_vslow_host, c->h_vslow, 0) != hipSuccess ||
	   !c->d_vslow_host)
		return bail(fail(PP_ERR_NOMEM, "pinned host flag allocation failed"));
	c->h_vslow[0] = 0;
	c->h_vslow[1] = 0;
	// diagnostic switches for the box trials' energy, which must give the
	// same trajectories: PP_NPT_FULL_TRIAL=1 regroups the particles into the
	// spare buffer for every trial; PP_NPT_PARTICLE_TRIAL=1 evaluates fast
	// trials per particle instead of from the cell records;
	// PP_NPT_RECORD_TRIAL=1 takes them from the records, never from the
	// close-pair list
	{
		auto on = [](const char *name) {
			const char *e = getenv(name);
			return e && *e && strcmp(e, "0") != 0;
		};
		d.npt_trial_form = on("PP_NPT_FULL_TRIAL") ? 2 : (on("PP_NPT_PARTICLE_TRIAL") ? 1 : (on("PP_NPT_RECORD_TRIAL") ? 3 : 0));
		// PP_NPT_STEPWISE=1: the host reads every box attempt's trial paths
		// before enqueueing its energy (no run-ahead; the same trajectories)
		c->vol_run_ahead = d.npt_trial_form == 0 && !on("PP_NPT_STEPWISE");
		// PP_SMALL_CELL_CAP=k (1 to the small-cell pass's capacity): defer
		// cells of more than k members to the general pass (a diagnostic
		// that must give the same trajectories)
		d.small_cap = small_cell_cap();
		if(const char *e = getenv("PP_SMALL_CELL_CAP"))
			if(*e) d.small_cap = std::max(1, std::min(small_cell_cap(), atoi(e)));
	}
	if(hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, o.device) != hipSuccess || c->cus <= 0)
		c->cus = 256;
	{
		// every buffer a kernel may touch must exist before the first launch
		const void *need[] = {d.f[0], d.f[1], d.pid[0], d.pid[1], d.reps, d.prep, d.act_off, d.count, d.start,
		                      d.slot, d.cellid, d.scan_status, d.tile_sum, d.clist, d.ccount, d.pfx, d.rec, d.scan_ticket,
		                      d.ctr, d.err, d.out_bonds, d.out_ovl, d.bar, d.vflag, d.vslow, d.done, d.ddq, d.ccheck,
		                      d.ovf, d.ovf_n};
		for(const void *q : need)
			if(!q) return bail(fail(PP_ERR_NOMEM, "a device buffer was not allocated"));
	}

	// host staging: SoA in particle-index order, slot j = particle j; the
	// free slots of grand-canonical replicas hold no particle (index -1)
	{
		std::vector<double> h((size_t) NCOMP * d.stride, 0.);
		std::vector<int> hp(d.stride, -1), hr(d.stride, 0);
		long long src = 0;  // the caller's arrays hold the replicas' particles back to back
		for(int k = 0; k < R; k++) {
			for(int i = 0; i < cap[k]; i++) hr[c->p_off[k] + i] = k;
			for(int i = 0; i < n[k]; i++, src++) {
				long long j = c->p_off[k] + i;
				for(int a = 0; a < 3; a++) h[(size_t) a * d.stride + j] = r[3 * src + a];
				for(int a = 0; a < 9; a++) h[(size_t) (3 + a) * d.stride + j] = O[9 * src + a];
				hp[j] = i;
			}
		}
		if(hipMemcpy(d.f[0], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
		   hipMemcpy(d.pid[0], hp.data(), hp.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
		   hipMemcpy(d.pid[1], hp.data(), hp.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
		   hipMemcpy((void *) d.prep, hr.data(), hr.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
		   hipMemcpy((void *) d.act_off, c->act_off.data(), (R + 1) * sizeof(int), hipMemcpyHostToDevice) !=
		       hipSuccess ||
		   hipMemcpy(d.reps, c->reps.data(), R * sizeof(Rep), hipMemcpyHostToDevice) != hipSuccess)
			return bail(fail(PP_ERR_HIP, "initial upload failed"));
	}
	launch_fill_i32(d.count, 0, d.C_total + 1, c->stream);
	launch_fill_i32(d.err, 0, 1, c->stream);
	launch_fill_i32((int *) d.bar, 0, 2, c->stream);
	launch_fill_i32(d.vflag, 0, R, c->stream);
	launch_fill_i32(d.vslow, 0, 1, c->stream);
	launch_fill_i32(d.vslow + 1, -1, 1, c->stream);
	launch_fill_i32((int *) d.done, 0, d.C_total + 1, c->stream);
	launch_fill_i32(d.ddq, 0, 1, c->stream);
	launch_fill_i32(d.ovf_n, 0, 8, c->stream);
	launch_fill_u64(d.ccheck, 0ull, 2, c->stream);
	if(d.pl_n) launch_fill_i32(d.pl_n, 0, 2, c->stream);
	launch_fill_u64(d.ctr, 0ull, (long long) R * NCTR * NSLOT, c->stream);

	// initial energy (main.c:54-61): any overlap is an error
	if((st = energy_now(c, 0)) != PP_OK) return bail(st);
	if((st = check_err(c)) != PP_OK) return bail(st);
	std::vector<unsigned long long> h;
	if((st = pull_ctr(c, h)) != PP_OK) return bail(st);
	if((st = pull_reps(c)) != PP_OK) return bail(st);
	for(int k = 0; k < R; k++) {
		// an energy-only context takes any snapshot; a sampler refuses overlaps
		if(!energy_only && ctr_sum(h, k, C_OVL) > 0)
			return bail(fail(PP_ERR_OVERLAP, "replica %d: initial configuration contains an overlap", k));
		c->reps[k].bonds_base = ctr_sum(h, k, C_EBONDS) / 2;
	}
	if((st = push_reps(c)) != PP_OK) return bail(st);
	*out = c;
	return PP_OK;
}

}  // namespace

extern "C" {

pp_status pp_create(const pp_params *params, const pp_options *opts, int n, const double *r, const double *O,
                    const double box[3], pp_ctx **out) {
	return create_impl(1, params, opts, &n, r, O, box, false, out);
}

pp_status pp_create_batch(int n_rep, const pp_params *params, const pp_options *opts, const int *n, const double *r,
                          const double *O, const double *box, pp_ctx **out) {
	return create_impl(n_rep, params, opts, n, r, O, box, false, out);
}

}  // extern "C"

namespace {

// Whether colour pass k of sweep sw uses the learned proposal: a fixed
// function of (seed, sweep, pass), true with probability `frac`, so the
// choice is independent of the state (a mixture of two kernels that each
// keep the Boltzmann distribution) and trajectories stay reproducible.
bool learned_pass(uint64_t seed, uint64_t sw, int k, double frac) {
	if(frac <= 0.) return false;
	if(frac >= 1.) return true;
	uint64_t z = seed ^ 0x6C8E9CF570932BD5ull;
	z += (sw * 8 + (uint64_t) k + 1) * 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	z ^= z >> 31;
	return (double) (z >> 11) * (1.0 / 9007199254740992.0) < frac;
}

// The live checkerboard checks (pp_colour_check): every replica on the grid
// of its current lists, added to the running total.
void colour_check_now(pp_ctx *c) {
	if(!c->colour_check) return;
	for(int k = 0; k < c->R; k++) launch_colour_conflicts(c->d, k, c->np[k], nullptr, c->d.ccheck, c->stream);
	c->colour_checks += c->R;
}

// Wait for the device to publish `seq` in pinned word `w` (k_vol_publish:
// an attempt's trial paths in the low 4 bits; k_vol_report: where a
// run-ahead sequence stopped in the low 8): a spin on host memory, with the
// stream polled now and then so that a failed launch is reported instead of
// waited for.
pp_status wait_published(pp_ctx *c, int w, int bits, unsigned seq, int *payload) {
	for(long long spin = 0;; spin++) {
		// (after the first few thousand polls, about 0.1 ms, the thread
		// yields between polls: a long pass ahead of the attempt need not
		// hold a core)
		if(spin > 4096) std::this_thread::yield();
		const int v = __atomic_load_n(c->h_vslow + w, __ATOMIC_ACQUIRE);
		if(((unsigned) v >> bits) == seq) {
			*payload = v & ((1 << bits) - 1);
			return PP_OK;
		}
		if((spin & 1023) == 1023) {
			const hipError_t q = hipStreamQuery(c->stream);
			if(q != hipSuccess && q != hipErrorNotReady) return fail(PP_ERR_HIP, "HIP error: %s", hipGetErrorString(q));
			if(q == hipSuccess && ((unsigned) __atomic_load_n(c->h_vslow + w, __ATOMIC_ACQUIRE) >> bits) != seq)
				return fail(PP_ERR_HIP, "the box attempts' trial paths were not published");
		}
		__builtin_ia32_pause();
	}
}

// Enqueue sweep c->sweep on the context's stream (no synchronisation).
pp_status enqueue_sweep(pp_ctx *c) {
	const uint64_t sw = (uint64_t) c->sweep;
	c->timing = c->timing_every > 0 && c->sweep % c->timing_every == 0;
	hipEvent_t a{};
	tic(c, TK_REBUILD, &a);
	long long tp = -1;
	if(c->trace_len >= 0 && c->trace_len < c->d.trace_cap) tp = c->trace_len++;
	launch_rebuild(c->d, REB_SHIFT, sw, tp, c->stream);
	toc(c, TK_REBUILD, a);
	colour_check_now(c);
	if(c->coop_blocks > 0 && !c->net) {
		hipEvent_t b{};
		tic(c, TK_PASS, &b);
		HIPCHECK(launch_sweep_passes(c->d, sw, c->T, c->block, c->coop_blocks, c->stream));
		toc(c, TK_PASS, b);
		colour_check_now(c);
	}
	else if(c->dd && !c->net) {
		// a new mark per sweep; when the marks wrap, the old ones are cleared
		if(++c->dd_mark == 0u) {
			launch_fill_i32((int *) c->d.done, 0, c->d.C_total + 1, c->stream);
			c->dd_mark = 1u;
		}
		hipEvent_t b{};
		tic(c, TK_PASS, &b);
		launch_passes_dd(c->d, sw, c->dd_mark, c->T, c->block, c->stream);
		toc(c, TK_PASS, b);
		colour_check_now(c);
	}
	else
		for(int k = 0; k < 8; k++) {
			hipEvent_t b{};
			if(c->net && learned_pass(c->d.seed, sw, k, c->learned_fraction)) {
				tic(c, TK_LEARN, &b);
				launch_learned_pass(c->d, c->net->d, sw, k, c->stream);
				toc(c, TK_LEARN, b);
			}
			else {
				tic(c, TK_PASS, &b);
				launch_pass(c->d, sw, k, c->T, c->block, c->stream);
				toc(c, TK_PASS, b);
			}
			colour_check_now(c);
		}
	// grand-canonical insertions and removals (MC_add_remove, MC.c:342-448)
	// on the lists the passes leave; box-shape trials after them rebuild
	if(c->any_gc) launch_gc_phase(c->d, sw, c->stream);
	if(c->any_npt || c->any_lx) {
		// one volume (or box-shape) attempt per N particle attempts on
		// average, as do_NPT and do_NVT (MC.c:19-29, 60-68): with s trials per
		// non-empty cell a step makes about s x (non-empty cells) <= s N
		// attempts, and s volume attempts
		const int nvol = c->d.trials_per_cell > 0 ? c->d.trials_per_cell : 1;
		// the close pairs of this sweep's state, for the trials it serves
		if(c->d.pl) launch_pairs_build(c->d, c->cus, c->stream);
		int k0 = 0;
		bool resume = false;
		if(c->vol_run_ahead && c->d.pl) {
			// every attempt enqueued on the close-pair list, which serves
			// every fast trial in the common case; the host then learns once
			// where, if anywhere, an attempt needed more (k_vol_gate), and
			// goes on from that attempt's standing proposal as below
			for(int k = 0; k < nvol; k++) {
				hipEvent_t v{};
				tic(c, TK_VOLUME, &v);
				launch_vol_propose(c->d, sw, k, c->stream);
				launch_vol_gate(c->d, k, c->stream);
				launch_energy_rows_zero(c->d, c->stream);
				launch_energy_pairs(c->d, c->cus, c->stream);
				launch_vol_decide(c->d, sw, k, c->stream);
				launch_vol_apply(c->d, c->stream);
				toc(c, TK_VOLUME, v);
			}
			const unsigned seq = (++c->vol_seq) & 0x7FFFFFu;
			launch_vol_report(c->d, c->d_vslow_host + 1, seq, c->stream);
			int stop = 0;
			TRY(wait_published(c, 1, 8, seq, &stop));
			k0 = stop > 0 ? stop - 1 : nvol;
			resume = stop > 0;
		}
		for(int k = k0; k < nvol; k++) {
			hipEvent_t v{};
			tic(c, TK_VOLUME, &v);
			if(!(resume && k == k0)) launch_vol_propose(c->d, sw, k, c->stream);
			// a trial rebuild only when some replica's trial box cannot be
			// served by its current lists (MC.c:481-485 rebuilds only when
			// the cells become too small); then the trial energies, per
			// particle and from the cell records, as the trials need them
			const unsigned seq = (++c->vol_seq) & 0x7FFFFFFu;
			launch_vol_publish(c->d, c->d_vslow_host, seq, c->stream);
			int need = 0;
			TRY(wait_published(c, 0, 4, seq, &need));
			if(need & NEED_REBUILD) launch_rebuild(c->d, REB_TRIAL, sw, -1, c->stream);
			if(need & NEED_PARTICLES) launch_energy(c->d, 1, 0, c->stream);
			else launch_energy_rows_zero(c->d, c->stream);
			if(need & NEED_RECORDS) launch_energy_rec(c->d, c->cus, c->stream);
			if(need & NEED_PAIRS) launch_energy_pairs(c->d, c->cus, c->stream);
			launch_vol_decide(c->d, sw, k, c->stream);
			launch_vol_apply(c->d, c->stream);
			toc(c, TK_VOLUME, v);
		}
	}
	c->sweep++;
	if(c->timing && c->pending.size() > 4096) TRY(collect_timing(c));
	return PP_OK;
}

}  // namespace

extern "C" {

pp_status pp_sweeps(pp_ctx *c, long long nsw) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(nsw < 0) return fail(PP_ERR_ARG, "negative sweep count");
	HIPCHECK(hipSetDevice(c->device));
	for(long long s = 0; s < nsw; s++) TRY(enqueue_sweep(c));
	c->timing = c->timing_every > 0;
	TRY(collect_timing(c));
	return check_err(c);
}

pp_status pp_autotune(pp_ctx *c, int sweeps, int *threads_per_cell, int *block_size, double *us_per_sweep) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(sweeps < 1) return fail(PP_ERR_ARG, "sweeps per candidate must be >= 1");
	if(c->coop_blocks > 0 || c->dd)
		return fail(PP_ERR_UNSUPPORTED, "the autotuner works on the per-pass launch form");
	HIPCHECK(hipSetDevice(c->device));
	TRY(collect_timing(c));
	const int saved_every = c->timing_every;
	c->timing_every = 0;
	hipEvent_t e0, e1;
	HIPCHECK(hipEventCreate(&e0));
	HIPCHECK(hipEventCreate(&e1));
	int best_T = c->T, best_block = c->block;
	float best = -1.f;
	pp_status st = PP_OK;
	// every candidate runs real sweeps: the trajectory does not depend on the
	// launch shape, so nothing is wasted
	const int Ts[3] = {16, 32, 64}, blocks[3] = {64, 128, 256};
	for(int ti = 0; ti < 3 && st == PP_OK; ti++)
		for(int bi = 0; bi < 3 && st == PP_OK; bi++) {
			if(pass_lds_bytes(Ts[ti], blocks[bi]) > MAX_LDS_BYTES || blocks[bi] < Ts[ti]) continue;
			c->T = Ts[ti];
			c->block = blocks[bi];
			st = enqueue_sweep(c);  // warm-up of this shape
			if(st != PP_OK) break;
			if(hipEventRecord(e0, c->stream) != hipSuccess) {
				st = fail(PP_ERR_HIP, "hipEventRecord failed");
				break;
			}
			for(int k = 0; k < sweeps && st == PP_OK; k++) st = enqueue_sweep(c);
			if(st != PP_OK) break;
			float ms = 0.f;
			if(hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
			   hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
				st = fail(PP_ERR_HIP, "timing a candidate failed");
				break;
			}
			if(best < 0.f || ms < best) {
				best = ms;
				best_T = c->T;
				best_block = c->block;
			}
		}
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	c->T = best_T;
	c->block = best_block;
	c->timing_every = saved_every;
	c->timing = saved_every > 0;
	TRY(st);
	if(threads_per_cell) *threads_per_cell = best_T;
	if(block_size) *block_size = best_block;
	if(us_per_sweep) *us_per_sweep = best * 1e3 / sweeps;
	return check_err(c);
}

pp_status pp_energy(pp_ctx *c, int rep, int32_t *bonds, uint8_t *overlap, long long *total_bonds, int *any_overlap) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	HIPCHECK(hipSetDevice(c->device));
	TRY(energy_now(c, 1));
	TRY(collect_timing(c));
	TRY(check_err(c));
	TRY(pull_reps(c));
	const Rep &rp = c->reps[rep];
	const int n = rp.np, off = rp.p_off;
	if(!rp.gc) {
		if(bonds && n) HIPCHECK(hipMemcpy(bonds, c->d.out_bonds + off, n * sizeof(int32_t), hipMemcpyDeviceToHost));
		if(overlap && n) HIPCHECK(hipMemcpy(overlap, c->d.out_ovl + off, n * sizeof(uint8_t), hipMemcpyDeviceToHost));
	}
	else if(n) {
		// the results sit at their particles' indices: gather them in index order
		std::vector<int> ids, sorted;
		TRY(slot_indices(c, rep, ids, sorted));
		std::vector<int32_t> hb(rp.cap_np);
		std::vector<uint8_t> ho(rp.cap_np);
		HIPCHECK(hipMemcpy(hb.data(), c->d.out_bonds + off, rp.cap_np * sizeof(int32_t), hipMemcpyDeviceToHost));
		HIPCHECK(hipMemcpy(ho.data(), c->d.out_ovl + off, rp.cap_np * sizeof(uint8_t), hipMemcpyDeviceToHost));
		for(int i = 0; i < n; i++) {
			if(bonds) bonds[i] = hb[sorted[i]];
			if(overlap) overlap[i] = ho[sorted[i]];
		}
	}
	std::vector<unsigned long long> h;
	TRY(pull_ctr(c, h));
	if(total_bonds) *total_bonds = ctr_sum(h, rep, C_EBONDS) / 2;
	if(any_overlap) *any_overlap = ctr_sum(h, rep, C_OVL) > 0;
	return PP_OK;
}

pp_status pp_check_energy(pp_ctx *c, double *energy_per_rep) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	HIPCHECK(hipSetDevice(c->device));
	TRY(energy_now(c, 0));
	TRY(collect_timing(c));
	TRY(check_err(c));
	std::vector<unsigned long long> h;
	TRY(pull_ctr(c, h));
	TRY(pull_reps(c));
	// cells_check (cells.c:65-76): every particle sits in exactly one cell
	std::vector<int> st(2);
	for(int k = 0; k < c->R; k++) {
		const Rep &rp = c->reps[k];
		HIPCHECK(hipMemcpy(&st[0], c->d.start + rp.cell_off, sizeof(int), hipMemcpyDeviceToHost));
		HIPCHECK(hipMemcpy(&st[1], c->d.start + rp.cell_off + rp.cell_cap, sizeof(int), hipMemcpyDeviceToHost));
		if(st[1] - st[0] != rp.np)
			return fail(PP_ERR_CELLS, "replica %d: there are %d particles in cells, there should be %d", k,
			            st[1] - st[0], rp.np);
	}
	pp_status result = PP_OK;
	for(int k = 0; k < c->R; k++) {
		Rep &rp = c->reps[k];
		const long long db = ctr_sum(h, k, C_DBONDS);
		const double E_old = -(double) (rp.bonds_base + db);
		const long long B = ctr_sum(h, k, C_EBONDS) / 2;
		const double E = -(double) B;
		if(ctr_sum(h, k, C_OVL) > 0) {
			result = fail(PP_ERR_OVERLAP, "replica %d: computing energy from scratch resulted in an overlap", k);
			continue;
		}
		// MC.c:780-782
		if(std::fabs(E_old) > 1e-5 && std::fabs((E - E_old) / E_old) > 1e-5 && result == PP_OK)
			result = fail(PP_ERR_ENERGY, "replica %d: energy check failed, old energy = %lf, new energy = %lf", k, E_old,
			              E);
		rp.bonds_base = B - db;
		if(energy_per_rep) energy_per_rep[k] = E;
	}
	TRY(push_reps(c));
	launch_gram_schmidt(c->d, c->stream);
	HIPCHECK(hipStreamSynchronize(c->stream));
	return result;
}

pp_status pp_download(pp_ctx *c, int rep, double *r, double *O, double box[3]) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	HIPCHECK(hipSetDevice(c->device));
	if(c->any_gc) {
		// grand-canonical replicas: their live particles first (a rebuild)
		rebuild_now(c);
		HIPCHECK(hipStreamSynchronize(c->stream));
	}
	TRY(pull_reps(c));
	const Rep &rp = c->reps[rep];
	const int n = rp.np, off = rp.p_off, S = c->d.stride;
	std::vector<double> h((size_t) NCOMP * std::max(n, 1));
	std::vector<int> hp, sorted;
	for(int a = 0; a < NCOMP && n; a++)
		HIPCHECK(hipMemcpy(h.data() + (size_t) a * n, c->d.f[rp.sel] + (size_t) a * S + off, n * sizeof(double),
		                   hipMemcpyDeviceToHost));
	TRY(slot_indices(c, rep, hp, sorted));
	for(int j = 0; j < n; j++) {
		const int i = rp.gc ? rank_of(sorted, hp[j]) : hp[j];
		if(r)
			for(int a = 0; a < 3; a++) r[3 * i + a] = h[(size_t) a * n + j];
		if(O)
			for(int a = 0; a < 9; a++) O[9 * i + a] = h[(size_t) (3 + a) * n + j];
	}
	if(box)
		for(int a = 0; a < 3; a++) box[a] = rp.g.box[a];
	return PP_OK;
}

pp_status pp_snapshot_start(pp_ctx *c, int rep) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	if(c->snap_pending) return fail(PP_ERR_ARG, "a snapshot is still pending (pp_snapshot_finish)");
	HIPCHECK(hipSetDevice(c->device));
	if(!c->snap_cap) {
		// the staging area holds the largest replica; each piece is kept as it
		// is made (the destructor frees it), and snap_cap is set only once
		// all exist, so a failed first call is retried whole
		int cap = 1;
		for(int k = 0; k < c->R; k++) cap = std::max(cap, c->np[k]);
		if(!c->copy_stream && hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess) {
			c->copy_stream = nullptr;
			return fail(PP_ERR_HIP, "snapshot stream creation failed");
		}
		if(!c->snap_ready) HIPCHECK(hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming));
		if(!c->snap_done) HIPCHECK(hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming));
		if(!c->snap_dev) TRY(dev_alloc(c, &c->snap_dev, (size_t) NCOMP * cap));
		if(!c->snap_dev_pid) TRY(dev_alloc(c, &c->snap_dev_pid, (size_t) cap));
		if(!c->snap_host &&
		   hipHostMalloc((void **) &c->snap_host, sizeof(double) * NCOMP * cap, hipHostMallocDefault) != hipSuccess)
			c->snap_host = nullptr;
		if(!c->snap_host_pid &&
		   hipHostMalloc((void **) &c->snap_host_pid, sizeof(int) * cap, hipHostMallocDefault) != hipSuccess)
			c->snap_host_pid = nullptr;
		if(!c->snap_host || !c->snap_host_pid) return fail(PP_ERR_NOMEM, "pinned snapshot buffers");
		c->snap_cap = cap;
	}
	if(c->any_gc) {
		// grand-canonical replicas: their live particles first (a rebuild)
		rebuild_now(c);
	}
	TRY(pull_reps(c));
	const Rep &rp = c->reps[rep];
	const int n = rp.np, off = rp.p_off, S = c->d.stride, cap = c->snap_cap;
	// on the sweeps' stream: the slots into the staging area, then the copy
	// to the host on the snapshot stream once they are there
	for(int a = 0; a < NCOMP && n; a++)
		HIPCHECK(hipMemcpyAsync(c->snap_dev + (size_t) a * cap, c->d.f[rp.sel] + (size_t) a * S + off,
		                        sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
	if(n)
		HIPCHECK(hipMemcpyAsync(c->snap_dev_pid, c->d.pid[rp.sel] + off, sizeof(int) * n, hipMemcpyDeviceToDevice,
		                        c->stream));
	HIPCHECK(hipEventRecord(c->snap_ready, c->stream));
	HIPCHECK(hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
	if(n) {
		HIPCHECK(hipMemcpyAsync(c->snap_host, c->snap_dev, sizeof(double) * NCOMP * cap, hipMemcpyDeviceToHost,
		                        c->copy_stream));
		HIPCHECK(hipMemcpyAsync(c->snap_host_pid, c->snap_dev_pid, sizeof(int) * n, hipMemcpyDeviceToHost,
		                        c->copy_stream));
	}
	HIPCHECK(hipEventRecord(c->snap_done, c->copy_stream));
	c->snap_n = n;
	c->snap_gc = rp.gc != 0;
	for(int a = 0; a < 3; a++) c->snap_box[a] = rp.g.box[a];
	c->snap_pending = true;
	return PP_OK;
}

pp_status pp_snapshot_finish(pp_ctx *c, double *r, double *O, double box[3]) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(!c->snap_pending) return fail(PP_ERR_ARG, "no snapshot pending (pp_snapshot_start)");
	HIPCHECK(hipSetDevice(c->device));
	c->snap_pending = false;
	HIPCHECK(hipEventSynchronize(c->snap_done));
	const int n = c->snap_n, cap = c->snap_cap;
	std::vector<int> sorted(c->snap_host_pid, c->snap_host_pid + n);
	std::sort(sorted.begin(), sorted.end());
	for(int i = 0; i < n; i++)
		if(sorted[i] < 0 || (i > 0 && sorted[i] == sorted[i - 1]) || (!c->snap_gc && sorted[i] != i))
			return fail(PP_ERR_HIP, "corrupt particle index %d", sorted[i]);
	for(int j = 0; j < n; j++) {
		const int id = c->snap_host_pid[j];
		const int i = c->snap_gc ? rank_of(sorted, id) : id;
		if(r)
			for(int a = 0; a < 3; a++) r[3 * i + a] = c->snap_host[(size_t) a * cap + j];
		if(O)
			for(int a = 0; a < 9; a++) O[9 * i + a] = c->snap_host[(size_t) (3 + a) * cap + j];
	}
	if(box)
		for(int a = 0; a < 3; a++) box[a] = c->snap_box[a];
	return PP_OK;
}

}  // extern "C"

namespace {

// Replace replica rep's particles by n given ones (n = its particle count,
// or for a grand-canonical replica any count up to its GC_N_max).
pp_status upload_impl(pp_ctx *c, int rep, int n, const double *r, const double *O, const double box[3]) {
	HIPCHECK(hipSetDevice(c->device));
	TRY(pull_reps(c));
	Rep &rp = c->reps[rep];
	if(n != rp.np && !(rp.gc && n >= 0 && n <= rp.cap_np))
		return fail(PP_ERR_ARG, "replica %d holds %d particles, not %d", rep, rp.np, n);
	const int off = rp.p_off, S = c->d.stride;
	if(n && (!r || !O)) return fail(PP_ERR_ARG, "positions/orientations are NULL");
	if(rp.gc) {
		// the replica's other slots are free
		std::vector<int> none(rp.cap_np, -1);
		for(int b = 0; b < 2; b++)
			HIPCHECK(hipMemcpy(c->d.pid[b] + off, none.data(), sizeof(int) * rp.cap_np, hipMemcpyHostToDevice));
		rp.np = n;
		rp.np_hw = n;
	}
	if(box) {
		int gn[3];
		for(int a = 0; a < 3; a++)
			if(!(box[a] > 0.)) return fail(PP_ERR_ARG, "box side %d is not positive", a);
		if(!make_grid(box, rp.side_min, rp.min_n, rp.even, rp.cap_n, gn))
			return fail(PP_ERR_BOX, "box too small for the cell grid");
		set_grid_box(rp.g, box);
		for(int a = 0; a < 3; a++) {
			rp.g.n[a] = gn[a];
			rp.g.shift[a] = 0.;
		}
		rp.V = box[0] * box[1] * box[2];
	}
	std::vector<double> h((size_t) NCOMP * std::max(n, 1));
	std::vector<int> hp(std::max(n, 1));
	for(int i = 0; i < n; i++) {
		for(int a = 0; a < 3; a++) h[(size_t) a * n + i] = r[3 * i + a];
		for(int a = 0; a < 9; a++) h[(size_t) (3 + a) * n + i] = O[9 * i + a];
		hp[i] = i;
	}
	for(int a = 0; a < NCOMP && n; a++)
		HIPCHECK(hipMemcpy(c->d.f[rp.sel] + (size_t) a * S + off, h.data() + (size_t) a * n, n * sizeof(double),
		                   hipMemcpyHostToDevice));
	if(n) HIPCHECK(hipMemcpy(c->d.pid[rp.sel] + off, hp.data(), n * sizeof(int), hipMemcpyHostToDevice));
	TRY(push_reps(c));
	// new tracked energy for this replica
	TRY(energy_now(c, 0));
	TRY(check_err(c));
	std::vector<unsigned long long> hc;
	TRY(pull_ctr(c, hc));
	TRY(pull_reps(c));
	if(ctr_sum(hc, rep, C_OVL) > 0) return fail(PP_ERR_OVERLAP, "uploaded configuration contains an overlap");
	c->reps[rep].bonds_base = ctr_sum(hc, rep, C_EBONDS) / 2 - ctr_sum(hc, rep, C_DBONDS);
	return push_reps(c);
}

}  // namespace

extern "C" {

pp_status pp_upload(pp_ctx *c, int rep, const double *r, const double *O, const double box[3]) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	HIPCHECK(hipSetDevice(c->device));
	TRY(pull_reps(c));
	return upload_impl(c, rep, c->reps[rep].np, r, O, box);
}

pp_status pp_upload_n(pp_ctx *c, int rep, int n, const double *r, const double *O, const double box[3]) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	return upload_impl(c, rep, n, r, O, box);
}

pp_status pp_get_counters(pp_ctx *c, int rep, pp_counters *out) {
	if(!c || !out) return fail(PP_ERR_ARG, "NULL argument");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	HIPCHECK(hipSetDevice(c->device));
	std::vector<unsigned long long> h;
	TRY(pull_ctr(c, h));
	TRY(pull_reps(c));
	const Rep &rp = c->reps[rep];
	out->tries[0] = ctr_sum(h, rep, C_TRY_ROTO);
	out->accepted[0] = ctr_sum(h, rep, C_ACC_ROTO);
	out->tries[1] = ctr_sum(h, rep, C_TRY_VOL);
	out->accepted[1] = ctr_sum(h, rep, C_ACC_VOL);
	out->bonds = rp.bonds_base + ctr_sum(h, rep, C_DBONDS);
	out->sweeps = c->sweep;
	for(int a = 0; a < 3; a++) out->box[a] = rp.g.box[a];
	out->volume = rp.V;
	out->gc_tries[0] = ctr_sum(h, rep, C_TRY_ADD);
	out->gc_tries[1] = ctr_sum(h, rep, C_TRY_REM);
	out->gc_accepted[0] = ctr_sum(h, rep, C_ACC_ADD);
	out->gc_accepted[1] = ctr_sum(h, rep, C_ACC_REM);
	out->n = rp.np;
	return PP_OK;
}

pp_status pp_get_counters_all(pp_ctx *c, pp_counters *out) {
	if(!c || !out) return fail(PP_ERR_ARG, "NULL argument");
	HIPCHECK(hipSetDevice(c->device));
	std::vector<unsigned long long> h;
	TRY(pull_ctr(c, h));
	TRY(pull_reps(c));
	for(int rep = 0; rep < c->R; rep++) {
		const Rep &rp = c->reps[rep];
		pp_counters &o = out[rep];
		o.tries[0] = ctr_sum(h, rep, C_TRY_ROTO);
		o.accepted[0] = ctr_sum(h, rep, C_ACC_ROTO);
		o.tries[1] = ctr_sum(h, rep, C_TRY_VOL);
		o.accepted[1] = ctr_sum(h, rep, C_ACC_VOL);
		o.bonds = rp.bonds_base + ctr_sum(h, rep, C_DBONDS);
		o.sweeps = c->sweep;
		for(int a = 0; a < 3; a++) o.box[a] = rp.g.box[a];
		o.volume = rp.V;
		o.gc_tries[0] = ctr_sum(h, rep, C_TRY_ADD);
		o.gc_tries[1] = ctr_sum(h, rep, C_TRY_REM);
		o.gc_accepted[0] = ctr_sum(h, rep, C_ACC_ADD);
		o.gc_accepted[1] = ctr_sum(h, rep, C_ACC_REM);
		o.n = rp.np;
	}
	return PP_OK;
}

pp_status pp_reset_counters(pp_ctx *c) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	HIPCHECK(hipSetDevice(c->device));
	std::vector<unsigned long long> h;
	TRY(pull_ctr(c, h));
	for(int k = 0; k < c->R; k++)
		for(int row : {(int) C_TRY_ROTO, (int) C_ACC_ROTO, (int) C_TRY_VOL, (int) C_ACC_VOL, (int) C_TRY_LEARN,
		               (int) C_ACC_LEARN, (int) C_TRY_ADD, (int) C_ACC_ADD, (int) C_TRY_REM, (int) C_ACC_REM})
			for(int s = 0; s < NSLOT; s++) h[((size_t) k * NCTR + row) * NSLOT + s] = 0;
	HIPCHECK(hipMemcpy(c->d.ctr, h.data(), h.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
	return PP_OK;
}

int pp_num_replicas(const pp_ctx *c) { return c ? c->R : 0; }

int pp_num_particles(const pp_ctx *c, int rep) { return (c && rep >= 0 && rep < c->R) ? c->np[rep] : -1; }

pp_status pp_energy_snapshot(int n, const double *r, const double *O, const double box[3], double kf_delta,
                             double kf_cosmax, int32_t *bonds, uint8_t *overlap, long long *total_bonds,
                             int *any_overlap) {
	pp_params p{};
	p.ensemble = PP_NVT;
	p.temperature = 1.;
	p.kf_delta = kf_delta;
	p.kf_cosmax = kf_cosmax;
	pp_ctx *c = nullptr;
	pp_status s = create_impl(1, &p, nullptr, &n, r, O, box, true, &c);
	if(s != PP_OK) return s;
	s = pp_energy(c, 0, bonds, overlap, total_bonds, any_overlap);
	pp_destroy(c);
	return s;
}

pp_status pp_set_trace(pp_ctx *c, int capacity) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(capacity < 0) return fail(PP_ERR_ARG, "negative capacity");
	HIPCHECK(hipSetDevice(c->device));
	HIPCHECK(hipStreamSynchronize(c->stream));
	if(c->d.trace) {
		HIPCHECK(hipFree(c->d.trace));
		c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void *) c->d.trace));
		c->d.trace = nullptr;
	}
	c->d.trace_cap = 0;
	c->trace_len = -1;
	if(capacity > 0) {
		TRY(dev_alloc(c, &c->d.trace, (size_t) c->R * capacity));
		c->d.trace_cap = capacity;
		c->trace_len = 0;
		c->trace_from = c->sweep;
	}
	return PP_OK;
}

pp_status pp_get_trace(pp_ctx *c, int rep, long long *bonds, int capacity, int *count) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	HIPCHECK(hipSetDevice(c->device));
	HIPCHECK(hipStreamSynchronize(c->stream));
	// entry k holds the energy after sweep k (written at the start of sweep k+1);
	// the state after the last sweep is appended from the counters
	int m = (int) std::max<long long>(c->trace_len, 0);
	int have = m > 0 ? m - 1 : 0;
	int nout = std::min(have, capacity);
	if(nout > 0 && bonds)
		HIPCHECK(hipMemcpy(bonds, c->d.trace + (size_t) rep * c->d.trace_cap + 1, nout * sizeof(long long),
		                   hipMemcpyDeviceToHost));
	// the energy after the last traced sweep is still in the counters only if
	// no sweep has run past the end of the trace
	if(bonds && nout < capacity && m > 0 && c->sweep - c->trace_from == m) {
		pp_counters k;
		TRY(pp_get_counters(c, rep, &k));
		bonds[nout++] = k.bonds;
	}
	if(count) *count = nout;
	return PP_OK;
}

pp_status pp_set_timing(pp_ctx *c, int every) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(every < 0) return fail(PP_ERR_ARG, "negative timing interval");
	c->timing_every = every;
	c->timing = every > 0;
	for(int k = 0; k < TK_KINDS; k++) {
		c->t_ms[k] = 0.;
		c->t_launch[k] = 0;
	}
	return PP_OK;
}

pp_status pp_get_timing(pp_ctx *c, const char *name, long long *launches, double *total_ms) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	TRY(collect_timing(c));
	int lo = 0, hi = TK_KINDS;
	if(name) {
		static const char *names[TK_KINDS] = {"pass", "rebuild", "energy", "volume", "learned"};
		lo = -1;
		for(int k = 0; k < TK_KINDS; k++)
			if(!strcmp(name, names[k])) lo = k;
		if(lo < 0) return fail(PP_ERR_ARG, "unknown kernel group '%s'", name);
		hi = lo + 1;
	}
	long long nl = 0;
	double ms = 0.;
	for(int k = lo; k < hi; k++) {
		nl += c->t_launch[k];
		ms += c->t_ms[k];
	}
	if(launches) *launches = nl;
	if(total_ms) *total_ms = ms;
	return PP_OK;
}

namespace {

// checkpoint file: header, then per replica its size, parameters, box,
// tracked volume, move counters, positions and orientations (FP64, bytes as
// held on the device)
constexpr char CHK_MAGIC[8] = {'P', 'P', 'C', 'H', 'K', '1', 0, 0};
constexpr int CHK_NROWS = 10;
constexpr int CHK_ROWS[CHK_NROWS] = {C_TRY_ROTO, C_ACC_ROTO, C_TRY_VOL, C_ACC_VOL, C_TRY_LEARN,
                                     C_ACC_LEARN, C_TRY_ADD, C_ACC_ADD, C_TRY_REM, C_ACC_REM};
constexpr int CHK_VERSION = 2;

struct ChkHead {
	char magic[8];
	int32_t version, R;
	int64_t sweep;
	uint64_t seed;
	int32_t trials_per_cell, pad;
};

struct ChkRep {
	int32_t n, ensemble;
	double T, disp_max, theta_max, rc2, cosmax, P, rfm;
	double box[3], V;
	int64_t ctr[CHK_NROWS];
};

}  // namespace

pp_status pp_checkpoint_save(pp_ctx *c, const char *path) {
	if(!c || !path) return fail(PP_ERR_ARG, "NULL argument");
	HIPCHECK(hipSetDevice(c->device));
	TRY(check_err(c));
	TRY(pull_reps(c));
	std::vector<unsigned long long> hc;
	TRY(pull_ctr(c, hc));
	const std::string tmp = std::string(path) + ".tmp";
	FILE *f = fopen(tmp.c_str(), "wb");
	if(!f) return fail(PP_ERR_ARG, "cannot open '%s' for writing", tmp.c_str());
	ChkHead h{};
	std::memcpy(h.magic, CHK_MAGIC, 8);
	h.version = CHK_VERSION;
	h.R = c->R;
	h.sweep = c->sweep;
	h.seed = c->d.seed;
	h.trials_per_cell = c->d.trials_per_cell;
	bool ok = fwrite(&h, sizeof h, 1, f) == 1;
	std::vector<double> r, O;
	for(int k = 0; k < c->R && ok; k++) {
		const Rep &rp = c->reps[k];
		ChkRep q{};
		q.n = rp.np;
		q.ensemble = rp.ensemble;
		q.T = rp.T;
		q.disp_max = rp.disp_max;
		q.theta_max = rp.theta_max;
		q.rc2 = rp.rc2;
		q.cosmax = rp.cosmax;
		q.P = rp.P;
		q.rfm = rp.rfm;
		for(int a = 0; a < 3; a++) q.box[a] = rp.g.box[a];
		q.V = rp.V;
		for(int t = 0; t < CHK_NROWS; t++) q.ctr[t] = ctr_sum(hc, k, CHK_ROWS[t]);
		r.assign(3 * (size_t) std::max(rp.np, 1), 0.);
		O.assign(9 * (size_t) std::max(rp.np, 1), 0.);
		pp_status st = pp_download(c, k, r.data(), O.data(), nullptr);
		if(st != PP_OK) {
			fclose(f);
			return st;
		}
		ok = fwrite(&q, sizeof q, 1, f) == 1 && fwrite(r.data(), sizeof(double), 3 * (size_t) rp.np, f) == 3 * (size_t) rp.np &&
		     fwrite(O.data(), sizeof(double), 9 * (size_t) rp.np, f) == 9 * (size_t) rp.np;
	}
	ok = (fclose(f) == 0) && ok;
	if(!ok || std::rename(tmp.c_str(), path) != 0) return fail(PP_ERR_ARG, "writing '%s' failed", path);
	return PP_OK;
}

pp_status pp_checkpoint_load(pp_ctx *c, const char *path) {
	if(!c || !path) return fail(PP_ERR_ARG, "NULL argument");
	HIPCHECK(hipSetDevice(c->device));
	FILE *f = fopen(path, "rb");
	if(!f) return fail(PP_ERR_ARG, "cannot open '%s'", path);
	ChkHead h{};
	std::vector<ChkRep> q(c->R);
	std::vector<std::vector<double>> r(c->R), O(c->R);
	pp_status st = PP_OK;
	if(fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, CHK_MAGIC, 8) != 0 || h.version != CHK_VERSION)
		st = fail(PP_ERR_ARG, "'%s' is not a checkpoint of this library", path);
	else if(h.R != c->R)
		st = fail(PP_ERR_ARG, "checkpoint holds %d replicas, the context %d", h.R, c->R);
	else if(h.seed != c->d.seed || h.trials_per_cell != c->d.trials_per_cell)
		st = fail(PP_ERR_ARG, "checkpoint written with another seed or trials_per_cell");
	TRY(pull_reps(c));
	for(int k = 0; k < c->R && st == PP_OK; k++) {
		const Rep &rp = c->reps[k];
		ChkRep &x = q[k];
		if(fread(&x, sizeof x, 1, f) != 1) {
			st = fail(PP_ERR_ARG, "checkpoint '%s' is truncated", path);
			break;
		}
		const bool n_ok = rp.gc ? (x.n >= 0 && x.n <= rp.cap_np) : x.n == rp.np;
		if(!n_ok || x.ensemble != rp.ensemble || x.T != rp.T || x.disp_max != rp.disp_max ||
		   x.theta_max != rp.theta_max || x.rc2 != rp.rc2 || x.cosmax != rp.cosmax || x.P != rp.P || x.rfm != rp.rfm) {
			st = fail(PP_ERR_ARG, "replica %d of the checkpoint has other parameters or size than the context", k);
			break;
		}
		r[k].resize(3 * (size_t) x.n);
		O[k].resize(9 * (size_t) x.n);
		if(fread(r[k].data(), sizeof(double), r[k].size(), f) != r[k].size() ||
		   fread(O[k].data(), sizeof(double), O[k].size(), f) != O[k].size())
			st = fail(PP_ERR_ARG, "checkpoint '%s' is truncated", path);
	}
	fclose(f);
	if(st != PP_OK) return st;
	for(int k = 0; k < c->R; k++) TRY(upload_impl(c, k, q[k].n, r[k].data(), O[k].data(), q[k].box));
	TRY(pull_reps(c));
	for(int k = 0; k < c->R; k++) c->reps[k].V = q[k].V;
	TRY(push_reps(c));
	std::vector<unsigned long long> hc;
	TRY(pull_ctr(c, hc));
	for(int k = 0; k < c->R; k++)
		for(int t = 0; t < CHK_NROWS; t++)
			for(int sl = 0; sl < NSLOT; sl++)
				hc[((size_t) k * NCTR + CHK_ROWS[t]) * NSLOT + sl] = sl == 0 ? (unsigned long long) q[k].ctr[t] : 0ull;
	HIPCHECK(hipMemcpy(c->d.ctr, hc.data(), hc.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
	c->sweep = h.sweep;
	return PP_OK;
}

pp_status pp_set_lx_move(pp_ctx *c, int rep, double lx_change_max, double lyz_min, double lyz_max) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < -1 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	HIPCHECK(hipSetDevice(c->device));
	TRY(pull_reps(c));
	const int lo = rep < 0 ? 0 : rep, hi = rep < 0 ? c->R : rep + 1;
	for(int k = lo; k < hi; k++) {
		const Rep &rp = c->reps[k];
		if(lx_change_max <= 0.) continue;  // off
		if(rp.ensemble != PP_NVT) return fail(PP_ERR_UNSUPPORTED, "box-shape moves need the NVT ensemble (replica %d)", k);
		if(rp.g.box[1] != rp.g.box[2]) return fail(PP_ERR_ARG, "Lx_move = 1 requires Ly = Lz");
		if(!(lyz_min < rp.g.box[1])) return fail(PP_ERR_ARG, "Lyz_min should be smaller than box size");
		if(!(lyz_max > rp.g.box[1])) return fail(PP_ERR_ARG, "Lyz_max should be larger than box size");
	}
	for(int k = lo; k < hi; k++) {
		Rep &rp = c->reps[k];
		rp.lx_move = lx_change_max > 0. ? 1 : 0;
		rp.lx_change_max = lx_change_max;
		rp.lyz_min = lyz_min;
		rp.lyz_max = lyz_max;
	}
	c->any_lx = false;
	for(const Rep &rp : c->reps) c->any_lx = c->any_lx || rp.lx_move;
	if(c->any_lx && !c->d.pl) {
		// the close-pair list of the box trials (as created for NPT contexts)
		c->d.pl_cap = (int) std::min<long long>(8LL * c->d.stride, 0x7fffff00LL);
		TRY(dev_alloc(c, &c->d.pl, (size_t) c->d.pl_cap));
		TRY(dev_alloc(c, &c->d.pl_n, 2));
		launch_fill_i32(c->d.pl_n, 0, 2, c->stream);
	}
	return push_reps(c);
}

pp_status pp_set_lx_acceptance(pp_ctx *c, int rep, int rule) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < -1 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	if(rule != PP_LX_METROPOLIS && rule != PP_LX_REFERENCE) return fail(PP_ERR_ARG, "unknown box-shape acceptance rule %d", rule);
	HIPCHECK(hipSetDevice(c->device));
	TRY(pull_reps(c));
	const int lo = rep < 0 ? 0 : rep, hi = rep < 0 ? c->R : rep + 1;
	for(int k = lo; k < hi; k++) c->reps[k].lx_ref_rule = rule == PP_LX_REFERENCE ? 1 : 0;
	return push_reps(c);
}

pp_status pp_bond_list(pp_ctx *c, int rep, int max_per, int32_t *count, int32_t *partners) {
	if(!c || !count || (max_per > 0 && !partners) || max_per < 0) return fail(PP_ERR_ARG, "bad arguments");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	HIPCHECK(hipSetDevice(c->device));
	rebuild_now(c);  // every particle in the cell of its position
	HIPCHECK(hipStreamSynchronize(c->stream));
	TRY(pull_reps(c));
	const Rep &rp = c->reps[rep];
	const int n = rp.np, off = rp.p_off;
	// a grand-canonical replica's results sit at its particles' indices
	std::vector<int> ids, sorted;
	if(rp.gc) TRY(slot_indices(c, rep, ids, sorted));
	const int span = rp.gc ? rp.cap_np : n;
	int *dc = nullptr, *dp = nullptr;
	const size_t np_all = (size_t) std::max(c->d.Ntot, 1);
	if(hipMalloc(&dc, sizeof(int) * np_all) != hipSuccess ||
	   hipMalloc(&dp, sizeof(int) * np_all * (size_t) std::max(max_per, 1)) != hipSuccess) {
		if(dc) (void) hipFree(dc);
		return fail(PP_ERR_NOMEM, "hipMalloc of the bond list failed");
	}
	launch_bond_list(c->d, max_per, dc, dp, c->stream);
	hipError_t e = hipStreamSynchronize(c->stream);
	if(e == hipSuccess) e = hipGetLastError();
	std::vector<int32_t> hc((size_t) std::max(span, 1)), hp((size_t) std::max(span, 1) * std::max(max_per, 1));
	if(e == hipSuccess && span > 0) e = hipMemcpy(hc.data(), dc + off, sizeof(int) * span, hipMemcpyDeviceToHost);
	if(e == hipSuccess && span > 0 && max_per > 0)
		e = hipMemcpy(hp.data(), dp + (size_t) off * max_per, sizeof(int) * (size_t) span * max_per, hipMemcpyDeviceToHost);
	(void) hipFree(dc);
	(void) hipFree(dp);
	if(e != hipSuccess) return fail(PP_ERR_HIP, "bond list: %s", hipGetErrorString(e));
	for(int i = 0; i < n; i++) {
		const int at = rp.gc ? sorted[i] : i;
		count[i] = hc[at];
		for(int k = 0; k < max_per && k < hc[at]; k++) {
			const int q = hp[(size_t) at * max_per + k];
			partners[(size_t) i * max_per + k] = rp.gc ? rank_of(sorted, q) : q;
		}
	}
	return check_err(c);
}

pp_status pp_get_learn_counters(pp_ctx *c, int rep, long long *tries, long long *accepted) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	HIPCHECK(hipSetDevice(c->device));
	std::vector<unsigned long long> h;
	TRY(pull_ctr(c, h));
	if(tries) *tries = ctr_sum(h, rep, C_TRY_LEARN);
	if(accepted) *accepted = ctr_sum(h, rep, C_ACC_LEARN);
	return PP_OK;
}

pp_status pp_net_attach(pp_ctx *c, pp_net *net, double learned_fraction, double theta_max) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(net && c->any_gc) return fail(PP_ERR_UNSUPPORTED, "learned proposals are not available with grand-canonical replicas");
	if(!net) {
		c->net = nullptr;
		c->learned_fraction = 0.;
		c->d.learn.on = 0;
		return PP_OK;
	}
	if(net->device != c->device) return fail(PP_ERR_ARG, "network on device %d, context on device %d", net->device, c->device);
	if(!(learned_fraction >= 0. && learned_fraction <= 1.)) return fail(PP_ERR_ARG, "learned_fraction must lie in [0, 1]");
	if(!(theta_max > 0. && theta_max <= M_PI)) return fail(PP_ERR_ARG, "theta_max must lie in (0, pi]");
	HIPCHECK(hipSetDevice(c->device));
	LearnDev &L = c->d.learn;
	if(!L.xf) {
		const size_t m = (size_t) std::max(c->d.total_act, 1);
		TRY(dev_alloc(c, &L.xf, m * NN_IN_PAD));
		TRY(dev_alloc(c, &L.xr, m * NN_IN_PAD));
		TRY(dev_alloc(c, &L.yf, m * NN_OUT_PAD));
		TRY(dev_alloc(c, &L.yr, m * NN_OUT_PAD));
		TRY(dev_alloc(c, &L.prop, m * NN_PROP_W));
		TRY(dev_alloc(c, &L.slot, m));
		HIPCHECK(hipMemsetAsync(L.xf, 0, m * NN_IN_PAD * sizeof(__bf16), c->stream));
		HIPCHECK(hipMemsetAsync(L.xr, 0, m * NN_IN_PAD * sizeof(__bf16), c->stream));
		HIPCHECK(hipStreamSynchronize(c->stream));
	}
	L.on = 1;
	L.theta_max = theta_max;
	c->net = net;
	c->learned_fraction = learned_fraction;
	return PP_OK;
}

pp_status pp_net_train(pp_ctx *c, pp_net *net, int steps, int batch, double lr, double pos_noise, double theta_max,
                       int sweeps_between, double *loss_first, double *loss_last) {
	if(!c || !net) return fail(PP_ERR_ARG, "NULL argument");
	if(c->any_gc) return fail(PP_ERR_UNSUPPORTED, "learned proposals are not available with grand-canonical replicas");
	if(net->device != c->device) return fail(PP_ERR_ARG, "network on device %d, context on device %d", net->device, c->device);
	if(steps < 0 || batch < 1 || batch > (1 << 20)) return fail(PP_ERR_ARG, "steps >= 0 and batch in [1, 2^20] required");
	if(!(lr > 0.) || !(pos_noise > 0.) || !(theta_max > 0. && theta_max <= M_PI) || sweeps_between < 0)
		return fail(PP_ERR_ARG, "lr > 0, pos_noise > 0, theta_max in (0, pi] and sweeps_between >= 0 required");
	HIPCHECK(hipSetDevice(c->device));
	TRY((pp_status) net_reserve(net, batch));
	HIPCHECK(hipStreamSynchronize(net->stream));
	std::vector<float> hl(batch);
	auto mean_loss = [&](double *out) -> pp_status {
		HIPCHECK(hipMemcpyAsync(hl.data(), net->loss, sizeof(float) * batch, hipMemcpyDeviceToHost, c->stream));
		HIPCHECK(hipStreamSynchronize(c->stream));
		double sum = 0.;
		for(float v : hl) sum += v;
		*out = sum / batch;
		return PP_OK;
	};
	const uint64_t tseed = c->d.seed ^ 0x2545F4914F6CDD1Dull;
	for(int s = 0; s < steps; s++) {
		if(s > 0)
			for(int k = 0; k < sweeps_between; k++) TRY(enqueue_sweep(c));
		// the cell ranges of the current state (an NPT trial may have left
		// those of its trial grid)
		rebuild_now(c);
		launch_train_samples(c->d, tseed, (uint64_t) net->step, batch, pos_noise, theta_max, net->x, net->target,
		                     c->stream);
		launch_net_loss_grad(net->d, net->x, net->target, batch, net->h, net->y, net->dy, net->dh, net->loss, c->stream);
		if(s == 0 && loss_first) TRY(mean_loss(loss_first));
		if(s == steps - 1 && loss_last) TRY(mean_loss(loss_last));
		net->step++;
		launch_net_adam(net->d, net->step, (float) lr, c->stream);
	}
	HIPCHECK(hipStreamSynchronize(c->stream));
	TRY(collect_timing(c));
	return check_err(c);
}

pp_status pp_get_grid(pp_ctx *c, int rep, int n_side[3], double side[3]) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	TRY(pull_reps(c));
	const Rep &rp = c->reps[rep];
	for(int a = 0; a < 3; a++) {
		if(n_side) n_side[a] = rp.g.n[a];
		if(side) side[a] = rp.g.box[a] / rp.g.n[a];
	}
	return PP_OK;
}

pp_status pp_colour_conflicts(pp_ctx *c, int rep, const int nside[3], long long *count) {
	if(!c || !count) return fail(PP_ERR_ARG, "bad arguments");
	if(rep < 0 || rep >= c->R) return fail(PP_ERR_ARG, "replica %d out of range", rep);
	if(nside)
		for(int k = 0; k < 3; k++)
			if(nside[k] < 1) return fail(PP_ERR_ARG, "nside[%d] = %d: at least one cell per axis", k, nside[k]);
	HIPCHECK(hipSetDevice(c->device));
	launch_fill_u64(c->d.ccheck + 1, 0ull, 1, c->stream);
	launch_colour_conflicts(c->d, rep, c->np[rep], nside, c->d.ccheck + 1, c->stream);
	unsigned long long v = 0;
	HIPCHECK(hipMemcpyAsync(&v, c->d.ccheck + 1, sizeof(v), hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	*count = (long long) v;
	return check_err(c);
}

pp_status pp_colour_check(pp_ctx *c, int on, long long *total, long long *checks) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(on < -1 || on > 1) return fail(PP_ERR_ARG, "on must be -1, 0 or 1");
	HIPCHECK(hipSetDevice(c->device));
	unsigned long long v = 0;
	HIPCHECK(hipMemcpyAsync(&v, c->d.ccheck, sizeof(v), hipMemcpyDeviceToHost, c->stream));
	HIPCHECK(hipStreamSynchronize(c->stream));
	if(total) *total = (long long) v;
	if(checks) *checks = c->colour_checks;
	if(on == 1) {
		launch_fill_u64(c->d.ccheck, 0ull, 1, c->stream);
		HIPCHECK(hipStreamSynchronize(c->stream));
		c->colour_checks = 0;
	}
	if(on >= 0) c->colour_check = on == 1;
	return PP_OK;
}

pp_status pp_get_layout(pp_ctx *c, long long *cell_capacity, int *scan_tile, int *scan_tiles) {
	if(!c) return fail(PP_ERR_ARG, "ctx is NULL");
	if(cell_capacity) *cell_capacity = c->d.C_total;
	if(scan_tile) *scan_tile = c->d.scan_tile;
	if(scan_tiles) *scan_tiles = c->d.ntiles;
	return PP_OK;
}

void pp_destroy(pp_ctx *c) {
	if(!c) return;
	(void) hipSetDevice(c->device);
	if(c->stream) (void) hipStreamSynchronize(c->stream);
	delete c;
}

const char *pp_strerror(pp_status s) {
	switch(s) {
	case PP_OK: return "success";
	case PP_ERR_ARG: return "invalid argument";
	case PP_ERR_HIP: return "HIP runtime error";
	case PP_ERR_OVERLAP: return "configuration contains an overlap";
	case PP_ERR_BOX: return "box too small for the cell grid";
	case PP_ERR_CELL_CAP: return "too many particles in one cell";
	case PP_ERR_ENERGY: return "energy check failed";
	case PP_ERR_NOMEM: return "out of memory";
	case PP_ERR_UNSUPPORTED: return "unsupported ensemble or dynamics";
	case PP_ERR_CELLS: return "cell population differs from N";
	}
	return "unknown status";
}

const char *pp_last_error(void) { return g_err.c_str(); }

}  // extern "C"
