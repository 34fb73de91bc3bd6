// pp_noalloc.cpp -- pp_energy_snapshot_noalloc: the per-particle energy of a
// snapshot (as pp_energy_snapshot) without any allocation, for callers that
// cannot allocate (code under stream capture, serving loops, harnesses that
// forbid hipMalloc).
//
// Every device buffer of the call -- the particle arrays in both layouts, the
// cell counts and starts, the scan's status words, the counters and the
// outputs -- lives in one file-scope device array, `scratch`, of
// NA_BYTES bytes.  Each GPU the library runs on holds its own copy of the
// array until the process ends; host code reaches it through the address
// hipGetSymbolAddress gives for that device (kept per device), never
// through the array's name.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU race, and the array keeps the last call's
// contents, so everything that must start at zero is cleared by a fill
// kernel at each call.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "../../include/pp.h"
#include "pp_internal.h"

using namespace ppd;

pp_status pp_set_error(pp_status s, const char *msg);  // pp_api.cpp

namespace {

// The largest snapshot and cell grid the array holds.  A box whose grid of
// cells of side 1 + delta would have more cells gets larger cells (fewer
// cells, more candidates per cell: the result is the same).
constexpr int NA_MAX_N = PP_NOALLOC_MAX_N;
constexpr int NA_MAX_CELLS = 1 << 18;
// scan tiles of a grid of at most NA_MAX_CELLS cells and its parking cell:
// small tiles up to SCAN_SMALL_MAX_TILES of them, large ones beyond
constexpr int NA_MAX_TILES = SCAN_SMALL_MAX_TILES > (NA_MAX_CELLS + SCAN_TILE_LARGE) / SCAN_TILE_LARGE
                                 ? SCAN_SMALL_MAX_TILES
                                 : (NA_MAX_CELLS + SCAN_TILE_LARGE) / SCAN_TILE_LARGE;

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t) 255; }

// Byte offsets of the buffers inside the array, each 256-byte aligned.
enum Part : int {
	P_F0 = 0, P_F1, P_PID0, P_PID1, P_PREP, P_SLOT, P_CELLID, P_BONDS, P_OVL, P_COUNT, P_START, P_STATUS, P_TILESUM,
	P_TICKET, P_CCOUNT, P_CTR, P_ERR, P_REPS, P_ACTOFF, NPART
};

constexpr size_t part_bytes(int p) {
	switch(p) {
	case P_F0:
	case P_F1: return sizeof(double) * NCOMP * (size_t) NA_MAX_N;
	case P_PID0:
	case P_PID1:
	case P_PREP:
	case P_SLOT:
	case P_CELLID:
	case P_BONDS: return sizeof(int) * (size_t) NA_MAX_N;
	case P_OVL: return (size_t) NA_MAX_N;
	case P_COUNT:
	case P_START: return sizeof(int) * ((size_t) NA_MAX_CELLS + 2);  // the grid, the parking cell, the end
	case P_STATUS: return sizeof(unsigned long long) * (size_t) NA_MAX_TILES;
	case P_TILESUM: return sizeof(int) * (size_t) NA_MAX_TILES;
	case P_TICKET: return sizeof(int) * 2;
	case P_CCOUNT: return sizeof(int) * 8;
	case P_CTR: return sizeof(unsigned long long) * NCTR * NSLOT;
	case P_ERR: return sizeof(int);
	case P_REPS: return sizeof(Rep);
	case P_ACTOFF: return sizeof(int) * 2;
	}
	return 0;
}

constexpr size_t part_off(int p) { return p == 0 ? 0 : align256(part_off(p - 1) + part_bytes(p - 1)); }

constexpr size_t NA_BYTES = align256(part_off(NPART - 1) + part_bytes(NPART - 1));
static_assert(NA_BYTES < (32u << 20), "a modest array: under 32 MiB per GPU");

}  // namespace

// The allocation-free call's device memory (see the file comment).
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
unsigned char *g_na_ptr[NA_MAX_DEVICES];  // device address of `scratch` on each device

// The address of device dev's copy of the array, looked up once per device
// with that device current.
pp_status na_scratch(int dev, unsigned char **out) {
	if(dev < 0 || dev >= NA_MAX_DEVICES) return pp_set_error(PP_ERR_ARG, "device ordinal out of range");
	std::lock_guard<std::mutex> lock(g_na_mutex);
	if(!g_na_ptr[dev]) {
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess || !p) {
			char buf[256];
			snprintf(buf, sizeof buf, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
			return pp_set_error(PP_ERR_HIP, buf);
		}
		g_na_ptr[dev] = (unsigned char *) p;
	}
	*out = g_na_ptr[dev];
	return PP_OK;
}

// Particles from the caller's layout (r[3j + k], O[9j + k], copied as they
// are into the second particle buffer) into the kernels' SoA layout in the
// first buffer, slot j = particle j of replica 0.
__global__ void k_na_stage(const double *aos, int n, double *f0, int *pid0, int *prep) {
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if(j >= n) return;
	for(int c = 0; c < 3; c++) f0[(size_t) c * n + j] = aos[3 * (size_t) j + c];
	const double *O = aos + 3 * (size_t) n;
	for(int c = 0; c < 9; c++) f0[(size_t) (3 + c) * n + j] = O[9 * (size_t) j + c];
	pid0[j] = j;
	prep[j] = 0;
}

#define NA_HIP(call)                                                                                      \
	do {                                                                                                  \
		const hipError_t e_ = (call);                                                                     \
		if(e_ != hipSuccess) {                                                                            \
			char b_[256];                                                                                 \
			snprintf(b_, sizeof b_, "%s failed: %s", #call, hipGetErrorString(e_));                       \
			return pp_set_error(PP_ERR_HIP, b_);                                                          \
		}                                                                                                 \
	} while(0)

}  // namespace

extern "C" pp_status pp_energy_snapshot_noalloc(int n, const double *r, const double *O, const double box[3],
                                                double kf_delta, double kf_cosmax, int32_t *bonds, uint8_t *overlap,
                                                long long *total_bonds, int *any_overlap) {
	char msg[256];
	if(n < 0 || n > NA_MAX_N) {
		snprintf(msg, sizeof msg, "n = %d: the allocation-free energy call takes 0 to %d particles", n, NA_MAX_N);
		return pp_set_error(PP_ERR_ARG, msg);
	}
	if(!box || (n > 0 && (!r || !O))) return pp_set_error(PP_ERR_ARG, "positions, orientations or box are NULL");
	if(!(kf_delta > 0.)) return pp_set_error(PP_ERR_ARG, "KF_delta must be > 0");
	for(int a = 0; a < 3; a++)
		if(!(box[a] > 0.)) return pp_set_error(PP_ERR_ARG, "box side is not positive");

	// the grid pp_energy_snapshot takes (at least 3 cells of side >= 1 + delta
	// per axis), with larger cells where it would exceed NA_MAX_CELLS
	const double rcut = 1. + kf_delta;
	const double V = box[0] * box[1] * box[2];
	const double side = std::fmax(rcut * (1. + 1e-9), std::cbrt(V / NA_MAX_CELLS) * (1. + 1e-12));
	int gn[3];
	if(!make_grid(box, side, 3, 0, nullptr, gn)) {
		snprintf(msg, sizeof msg, "box (%g, %g, %g) holds fewer than 3 cells of side %g per axis", box[0], box[1], box[2],
		         side);
		return pp_set_error(PP_ERR_BOX, msg);
	}
	const long long cells = (long long) gn[0] * gn[1] * gn[2];
	if(cells > NA_MAX_CELLS) return pp_set_error(PP_ERR_BOX, "cell grid beyond the allocation-free call's limit");
	if(n == 0) {
		if(total_bonds) *total_bonds = 0;
		if(any_overlap) *any_overlap = 0;
		return PP_OK;
	}

	NA_HIP(hipSetDevice(0));  // the device pp_energy_snapshot runs on
	unsigned char *base = nullptr;
	const pp_status st = na_scratch(0, &base);
	if(st != PP_OK) return st;
	auto at = [&](int p) { return (void *) (base + part_off(p)); };

	Rep rp{};
	rp.rc2 = rcut * rcut;  // SQR(1. + delta) (system.c:86)
	rp.cosmax = kf_cosmax;
	rp.T = 1.;
	for(int e = 0; e < ACC_TAB; e++) rp.acc_tab[e] = std::exp(-(double) e);
	rp.ensemble = PP_NVT;
	rp.min_n = 3;
	rp.even = 0;
	rp.side_min = side;
	for(int a = 0; a < 3; a++) {
		rp.g.n[a] = gn[a];
		rp.cap_n[a] = gn[a];
		rp.g.shift[a] = 0.;
	}
	set_grid_box(rp.g, box);
	rp.tg = rp.g;
	rp.V = V;
	rp.p_off = 0;
	rp.np = n;
	rp.cap_np = n;
	rp.np_hw = n;
	rp.cell_off = 0;
	rp.cell_cap = (int) cells;
	rp.act_cap = (gn[0] / 2) * (gn[1] / 2) * (gn[2] / 2);
	rp.sel = 0;

	DevState d{};
	d.f[0] = (double *) at(P_F0);
	d.f[1] = (double *) at(P_F1);
	d.pid[0] = (int *) at(P_PID0);
	d.pid[1] = (int *) at(P_PID1);
	d.stride = n;
	d.Ntot = n;
	d.R = 1;
	d.reps = (Rep *) at(P_REPS);
	d.prep = (const int *) at(P_PREP);
	d.act_off = (const int *) at(P_ACTOFF);
	d.total_act = rp.act_cap;
	d.C_total = (int) cells + 1;  // the grid and the replica's (empty) parking cell
	d.count = (int *) at(P_COUNT);
	d.start = (int *) at(P_START);
	d.slot = (int *) at(P_SLOT);
	d.cellid = (int *) at(P_CELLID);
	d.scan_status = (unsigned long long *) at(P_STATUS);
	d.ccount = (int *) at(P_CCOUNT);
	d.scan_ticket = (int *) at(P_TICKET);
	d.tile_sum = (int *) at(P_TILESUM);
	d.scan_tile = (d.C_total + SCAN_TILE_SMALL - 1) / SCAN_TILE_SMALL <= SCAN_SMALL_MAX_TILES ? SCAN_TILE_SMALL
	                                                                                         : SCAN_TILE_LARGE;
	d.ntiles = (d.C_total + d.scan_tile - 1) / d.scan_tile;
	if(d.ntiles > NA_MAX_TILES) return pp_set_error(PP_ERR_BOX, "scan tiles beyond the allocation-free call's limit");
	d.ctr = (unsigned long long *) at(P_CTR);
	d.err = (int *) at(P_ERR);
	d.out_bonds = (int32_t *) at(P_BONDS);
	d.out_ovl = (uint8_t *) at(P_OVL);

	// the caller's arrays into the second buffer as they are, then the SoA
	// layout into the first; the table of the one replica; zeroed counts and
	// error word (the array holds the previous call's values)
	hipStream_t s = nullptr;
	double *aos = d.f[1];
	NA_HIP(hipMemcpy(aos, r, sizeof(double) * 3 * (size_t) n, hipMemcpyHostToDevice));
	NA_HIP(hipMemcpy(aos + 3 * (size_t) n, O, sizeof(double) * 9 * (size_t) n, hipMemcpyHostToDevice));
	NA_HIP(hipMemcpy(d.reps, &rp, sizeof(Rep), hipMemcpyHostToDevice));
	hipLaunchKernelGGL(k_na_stage, dim3((n + 255) / 256), dim3(256), 0, s, aos, n, d.f[0], d.pid[0], (int *) d.prep);
	launch_fill_i32(d.count, 0, d.C_total + 1, s);
	launch_fill_i32(d.err, 0, 1, s);
	// regroup by cell (k_count, k_scan, k_scatter; the scan switches the
	// replica to the second buffer), then MC_energy of every particle
	launch_rebuild(d, REB_KEEP, 0, -1, s);
	launch_energy(d, 0, 1, s);
	NA_HIP(hipGetLastError());
	NA_HIP(hipStreamSynchronize(s));

	int err = 0;
	unsigned long long ctr[NCTR * NSLOT];
	NA_HIP(hipMemcpy(&err, d.err, sizeof(int), hipMemcpyDeviceToHost));
	if(err) return pp_set_error(PP_ERR_HIP, "cell scan of the allocation-free energy call did not complete");
	NA_HIP(hipMemcpy(ctr, d.ctr, sizeof ctr, hipMemcpyDeviceToHost));
	if(bonds) NA_HIP(hipMemcpy(bonds, d.out_bonds, sizeof(int32_t) * (size_t) n, hipMemcpyDeviceToHost));
	if(overlap) NA_HIP(hipMemcpy(overlap, d.out_ovl, (size_t) n, hipMemcpyDeviceToHost));
	unsigned long long eb = 0, ov = 0;
	for(int k = 0; k < NSLOT; k++) {
		eb += ctr[C_EBONDS * NSLOT + k];
		ov += ctr[C_OVL * NSLOT + k];
	}
	if(total_bonds) *total_bonds = (long long) (eb / 2);
	if(any_overlap) *any_overlap = ov > 0;
	return PP_OK;
}
