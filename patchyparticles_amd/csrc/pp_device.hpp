// pp_device.hpp -- device helpers shared by the gfx950 kernels of
// libpp_hip.so (pp_kernels.hip: rebuild, colour passes, energy; pp_learn.hip:
// the learned proposal): team operations, counters, trial draws and the
// reference-order pair tests on a trial.
#pragma once
#pragma clang fp contract(off)

#include "pp_internal.h"

namespace ppd {

// ------------------------------------------------------------------------
// helpers

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of logical blocks so that neighbouring cells share an L2.
// nb must be a multiple of 8.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
	int per = nb >> 3;
	return (b & 7) * per + (b >> 3);
}

// Buffer b (0 or 1) of the double-buffered slot arrays.  A team-uniform b
// that is not wavefront-uniform (a selector read from a cell record) would
// index the kernel argument with a vector load; later loads through the
// pointer then wait on every load issued before them.  Selecting between the
// two scalar base pointers keeps them in registers.
template <class P>
__device__ __forceinline__ P *pick_buf(P *const (&buf)[2], int b) {
	const unsigned long long p0 = (unsigned long long) buf[0], p1 = (unsigned long long) buf[1];
	const unsigned lo = b ? (unsigned) p1 : (unsigned) p0, hi = b ? (unsigned) (p1 >> 32) : (unsigned) (p0 >> 32);
	return (P *) (((unsigned long long) hi << 32) | lo);
}

// DPP controls (gfx9 encoding)
constexpr int DPP_QUAD_1032 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_2301 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int DPP_ROW_MIRROR = 0x140;    // lane i <- 15 - i within a 16-lane row
constexpr int DPP_HALF_MIRROR = 0x141;   // lane i <- 7 - i within 8 lanes

// Sum over a team of T = 8, 16, 32 or 64 lanes, in every lane of the team,
// with VALU cross-lane moves only (a __shfl_xor is a ds_bpermute, an LDS
// round trip per step): DPP butterflies within a 16-lane row, then gfx950's
// row-pair and half-wave swaps.  Each step stays inside the team, so lanes of
// the other teams of the wavefront may be inactive.
template <int T>
__device__ __forceinline__ int team_sum(int v) {
	static_assert(T == 8 || T == 16 || T == 32 || T == 64, "team of 8 to 64 lanes");
	v += __builtin_amdgcn_update_dpp(0, v, DPP_QUAD_1032, 0xf, 0xf, false);
	v += __builtin_amdgcn_update_dpp(0, v, DPP_QUAD_2301, 0xf, 0xf, false);
	v += __builtin_amdgcn_update_dpp(0, v, DPP_HALF_MIRROR, 0xf, 0xf, false);
	if constexpr(T >= 16) v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_MIRROR, 0xf, 0xf, false);
	if constexpr(T >= 32) {
		// rows 2j and 2j+1 trade places between two copies: copy sum = row-pair sum
		const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
		v = (int) (p[0] + p[1]);
	}
	if constexpr(T >= 64) {
		const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
		v = (int) (p[0] + p[1]);
	}
	return v;
}

// Sum of a double over a team of 16 lanes (DPP butterflies within a row, on
// the two halves of the value).  Lanes may round differently: for bounds
// only.
__device__ __forceinline__ double dpp_add_f64_quad1032(double x) {
	const long long b = __double_as_longlong(x);
	const int lo = __builtin_amdgcn_update_dpp(0, (int) b, DPP_QUAD_1032, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int) (b >> 32), DPP_QUAD_1032, 0xf, 0xf, false);
	return x + __longlong_as_double(((long long) hi << 32) | (unsigned) lo);
}
__device__ __forceinline__ double dpp_add_f64_quad2301(double x) {
	const long long b = __double_as_longlong(x);
	const int lo = __builtin_amdgcn_update_dpp(0, (int) b, DPP_QUAD_2301, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int) (b >> 32), DPP_QUAD_2301, 0xf, 0xf, false);
	return x + __longlong_as_double(((long long) hi << 32) | (unsigned) lo);
}
__device__ __forceinline__ double dpp_add_f64_half_mirror(double x) {
	const long long b = __double_as_longlong(x);
	const int lo = __builtin_amdgcn_update_dpp(0, (int) b, DPP_HALF_MIRROR, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int) (b >> 32), DPP_HALF_MIRROR, 0xf, 0xf, false);
	return x + __longlong_as_double(((long long) hi << 32) | (unsigned) lo);
}
__device__ __forceinline__ double dpp_add_f64_row_mirror(double x) {
	const long long b = __double_as_longlong(x);
	const int lo = __builtin_amdgcn_update_dpp(0, (int) b, DPP_ROW_MIRROR, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, (int) (b >> 32), DPP_ROW_MIRROR, 0xf, 0xf, false);
	return x + __longlong_as_double(((long long) hi << 32) | (unsigned) lo);
}
__device__ __forceinline__ double team16_sum_f64(double x) {
	x = dpp_add_f64_quad1032(x);
	x = dpp_add_f64_quad2301(x);
	x = dpp_add_f64_half_mirror(x);
	return dpp_add_f64_row_mirror(x);
}

// Lanes below this one whose bit is set in mask m.
__device__ __forceinline__ int lanes_below(unsigned long long m) {
	return (int) __builtin_amdgcn_mbcnt_hi((unsigned) (m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned) m, 0u));
}

// Inclusive prefix sum over a whole wavefront (every lane active):
// Hillis-Steele steps within each 16-lane row (row_shr 1, 2, 4, 8, the
// lanes shifted in from outside the row reading 0), then row 0's total into
// row 1 and row 2's into row 3 (row_bcast:15), then the first half's total
// into rows 2 and 3 (row_bcast:31).
constexpr int DPP_ROW_SHR = 0x110;     // + k: lane i <- i - k within a row
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

__device__ __forceinline__ int wave_incl_scan(int v) {
	v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR + 1, 0xf, 0xf, true);
	v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR + 2, 0xf, 0xf, true);
	v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR + 4, 0xf, 0xf, true);
	v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR + 8, 0xf, 0xf, true);
	v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST15, 0xa, 0xf, false);
	v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST31, 0xc, 0xf, false);
	return v;
}

// Sum over a team of small non-negative integers (< 8): one ballot per bit.
template <int T>
__device__ __forceinline__ int team_sum_small(int v) {
	const int off = (threadIdx.x & 63) & ~(T - 1);
	const unsigned long long tm = (T == 64) ? ~0ull : ((1ull << T) - 1ull);
	int s = 0;
#pragma unroll
	for(int b = 0; b < 3; b++) s += __popcll((__ballot((v >> b) & 1) >> off) & tm) << b;
	return s;
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

// Lanes of one team live in one wavefront: order their LDS traffic.
__device__ __forceinline__ void team_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned long long *ctr_at(unsigned long long *ctr, int r, int row, int slot) {
	return ctr + ((long long) (r * NCTR + row) * NSLOT + slot);
}

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

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

// Trial t of cell cloc (colour col) in sweep `sweep`: member rank among
// cnt, displacement uniform in +-disp_max/2 per axis (MC.c:317-319), rotation
// by an angle uniform in [0, theta_max) about an axis uniform on the sphere
// (MC.c:321-325, utils.c:186-202), and the acceptance draw (MC.c:332).
__device__ __forceinline__ void make_trial(const DevState &d, const Rep &rp, int r, uint64_t sweep, int cloc, int col,
                                           int t, int cnt, double *out) {
	u32x4 ra = rng(d.seed, r, sweep, (uint32_t) cloc, TAG_TRIAL_A, (uint32_t) col, (uint32_t) t);
	u32x4 rb = rng(d.seed, r, sweep, (uint32_t) cloc, TAG_TRIAL_B, (uint32_t) col, (uint32_t) t);
	out[0] = (double) __umulhi(ra.x, (uint32_t) cnt);
	out[1] = (u01(ra.y) - 0.5) * rp.disp_max;
	out[2] = (u01(ra.z) - 0.5) * rp.disp_max;
	out[3] = (u01(ra.w) - 0.5) * rp.disp_max;
	double axis[3], st, ct, R[9];
	sphere_axis(d.seed, r, sweep, (uint32_t) cloc, (uint32_t) col, (uint32_t) t, axis);
	sincos(u01(rb.x) * rp.theta_max, &st, &ct);
	rotation_matrix(axis, st, ct, R);
#pragma unroll
	for(int c = 0; c < 9; c++) out[4 + c] = R[c];
	out[13] = u01(rb.w);
}

// Trial geometry shared by the helpers of one team.
struct Trial {
	double px, py, pz, nx, ny, nz;
	double Lx, Ly, Lz, iLx, iLy, iLz, rc2, cosmax;
	double c2;  // cosmax^2 for the filtered angular tests, -1: exact tests only
};

// Minimum-image separation of q from the trial's new position, and its
// square (MC.c:232-240)
__device__ __forceinline__ double sep_new(const Trial &t, double qx, double qy, double qz, double &wx, double &wy,
                                          double &wz) {
	wx = kf_min_image(qx - t.nx, t.Lx, t.iLx);
	wy = kf_min_image(qy - t.ny, t.Ly, t.iLy);
	wz = kf_min_image(qz - t.nz, t.Lz, t.iLz);
	return wx * wx + wy * wy + wz * wz;
}

__device__ __forceinline__ bool overlaps_new(const Trial &t, double qx, double qy, double qz) {
	double dx, dy, dz;
	return sep_new(t, qx, qy, qz, dx, dy, dz) < 1.;
}

// Bonds of neighbour q with the old and the new position (one load of q for
// both: the particle stays in its cell, so the neighbourhood is the same),
// with the reference's arithmetic (MC.c:242-262).  q's orientation comes from
// load_o only when one of the two passes the p-side test.
template <class LoadO>
__device__ __forceinline__ void bonds_old_new_exact(const Trial &t, const double *Po, const double *Pn, double qx,
                                                    double qy, double qz, LoadO load_o, int &bo, int &bn) {
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
		po = kf_any_p(ox, oy, oz, Po, t.cosmax);
	}
	if(w2 < t.rc2) {
		kf_unit(wx, wy, wz, w2);
		pn = kf_any_p(wx, wy, wz, Pn, t.cosmax);
	}
	if(po || pn) {
		double Oq[9];
		load_o(Oq);
		bool ho, hn;
		kf_any_q2(ox, oy, oz, po, wx, wy, wz, pn, Oq, t.cosmax, ho, hn);
		bo += ho;
		bn += hn;
	}
}

// Bonds of neighbour q with the new position only: the filtered decision
// (a pair inside the filter band sets `band`) and the reference's arithmetic
// (MC.c:242-262).  For neighbours whose bonds with the old position are known.
// bonds_new_sep takes the separation (wx, wy, wz) = sep_new(q) and its
// square w2; bonds_new forms them.
template <class LoadO>
__device__ __forceinline__ void bonds_new_sep(const Trial &t, const double *Pn, double wx, double wy, double wz,
                                              double w2, LoadO load_o, int &bn, bool &band) {
	const KfBand bd = kf_band(t.c2, w2);
	const unsigned fn = w2 < t.rc2 ? kf_fast_p(wx, wy, wz, Pn, bd) : 0u;
	unsigned unsure = (fn >> 1) & 1u;
	if(fn & 1u) {
		double Oq[9];
		load_o(Oq);
		const unsigned qn = kf_fast_q1(wx, wy, wz, kf_band_q(bd), Oq);
		unsure |= (qn >> 1) & 1u;
		bn += (int) (qn & 1u);
	}
	band |= unsure != 0u;
}

template <class LoadO>
__device__ __forceinline__ void bonds_new(const Trial &t, const double *Pn, double qx, double qy, double qz,
                                          LoadO load_o, int &bn, bool &band) {
	double wx, wy, wz;
	const double w2 = sep_new(t, qx, qy, qz, wx, wy, wz);
	bonds_new_sep(t, Pn, wx, wy, wz, w2, load_o, bn, band);
}

template <class LoadO>
__device__ __forceinline__ void bonds_new_exact(const Trial &t, const double *Pn, double qx, double qy, double qz,
                                                LoadO load_o, int &bn) {
	double wx = kf_min_image(qx - t.nx, t.Lx, t.iLx);
	double wy = kf_min_image(qy - t.ny, t.Ly, t.iLy);
	double wz = kf_min_image(qz - t.nz, t.Lz, t.iLz);
	double w2 = wx * wx + wy * wy + wz * wz;
	if(!(w2 < t.rc2)) return;
	kf_unit(wx, wy, wz, w2);
	if(!kf_any_p(wx, wy, wz, Pn, t.cosmax)) return;
	double Oq[9];
	load_o(Oq);
	bn += kf_any_q(wx, wy, wz, Oq, t.cosmax);
}

}  // namespace ppd
