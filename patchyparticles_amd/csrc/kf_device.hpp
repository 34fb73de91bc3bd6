// kf_device.hpp -- device-side geometry of the Kern-Frenkel hot path for gfx950.
//
// Floating point: this header is compiled with -ffp-contract=off (and the
// pragma below), so every product and sum rounds where it is written.  The
// pair test reproduces the reference's operation order (MC.c:232-266,
// utils.h:14-21) so that bond and overlap classifications are bit-identical
// to the reference on the same FP64 snapshot.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ppd {

// Interaction classes (defs.h:31-33).
constexpr int KF_OVERLAP = -1;
constexpr int KF_NO_BOND = 0;
constexpr int KF_BOND = 1;

// Tetrahedral base patches (system.c:20-31): normalize() of (+-h,+-h,+-h)
// with h = 0.5/sqrt(3).  All four share the magnitude c = h / |(h,h,h)|,
// the double 0x1.279a74590331dp-1 (h = 0.2886751345948129, |.| = 0.5 exactly);
// the host recomputes it with the same IEEE operations at context creation
// and refuses to run if the bits differ.  Signs: patch i, component k is
// negative iff bit 3i+k of KF_NEG is set: (-,-,+) (+,-,-) (+,+,+) (-,+,-).
constexpr double KF_C = 0x1.279a74590331dp-1;
constexpr unsigned KF_NEG = 0b101000110011u;  // bits 0..11 = (i=0:k0,k1,k2), (i=1:...), ...

struct Base {
	double b[4][3];
};

__host__ __device__ constexpr double kf_base(int i, int k) { return ((KF_NEG >> (3 * i + k)) & 1u) ? -KF_C : KF_C; }

// patch_i = O * b_i over the rows of O, left-to-right sums (utils.h:17-21).
__device__ __forceinline__ void kf_patches(const double O[9], double P[4][3]) {
#pragma unroll
	for(int i = 0; i < 4; i++)
#pragma unroll
		for(int k = 0; k < 3; k++)
			P[i][k] = O[3 * k + 0] * kf_base(i, 0) + O[3 * k + 1] * kf_base(i, 1) + O[3 * k + 2] * kf_base(i, 2);
}

// Minimum-image separation d = r_q - r_p, component by component.  The image
// index comes from rint(d * (1/L)) instead of rint(d / L): the two agree
// whenever the component can matter (|d - kL| < 1+delta <= L/3 keeps d/L at
// least 1/6 away from a rounding boundary), and when they disagree the
// component is ~L/2 in both forms, so d^2 exceeds (1+delta)^2 either way.
// The subtraction d - L*k is then the reference's exact expression.
__device__ __forceinline__ double kf_min_image(double d, double L, double invL) {
	double k = __builtin_rint(d * invL);
	double t = L * k;
	return d - t;
}

__device__ __forceinline__ double kf_dot(double a0, double a1, double a2, double b0, double b1, double b2) {
	return a0 * b0 + a1 * b1 + a2 * b2;
}

// Angular part of MC_would_interact (MC.c:242-262): d = (dx,dy,dz) is the
// minimum-image r_q - r_p with rc2 > d2 >= 1.  Returns true on a bond.
__device__ __forceinline__ bool kf_bond(double dx, double dy, double dz, double d2, const double P[4][3],
                                        const double Q[4][3], double cosmax) {
	double len = __builtin_sqrt(d2);
	dx = dx / len;
	dy = dy / len;
	dz = dz / len;
	bool bond = false;
#pragma unroll
	for(int a = 0; a < 4; a++) {
		double cp = kf_dot(dx, dy, dz, P[a][0], P[a][1], P[a][2]);
		if(cp > cosmax) {
#pragma unroll
			for(int b = 0; b < 4; b++) {
				double cq = -kf_dot(dx, dy, dz, Q[b][0], Q[b][1], Q[b][2]);
				bond |= (cq > cosmax);
			}
		}
	}
	return bond;
}

// The same test split in its two halves.  MC.c:250-262 returns a bond as
// soon as some patch a of p has d^.p_a > cosmax and some patch b of q has
// -d^.q_b > cosmax (the inner condition does not depend on a), i.e.
// bond <=> any_p && any_q.  Evaluating any_q only when any_p holds lets the
// caller skip loading q's orientation for most pairs.
//
// d^ = d / |d|: one sqrt, three divisions (MC.c:244-248).
__device__ __forceinline__ void kf_unit(double &dx, double &dy, double &dz, double d2) {
	double len = __builtin_sqrt(d2);
	dx = dx / len;
	dy = dy / len;
	dz = dz / len;
}

// any patch of p (12 doubles, patch-major) with d^.p_a > cosmax
__device__ __forceinline__ bool kf_any_p(double ux, double uy, double uz, const double *P, double cosmax) {
	bool any = false;
#pragma unroll
	for(int a = 0; a < 4; a++) any |= kf_dot(ux, uy, uz, P[3 * a], P[3 * a + 1], P[3 * a + 2]) > cosmax;
	return any;
}

// any patch q_b = Oq * b_b of q with -(d^.q_b) > cosmax.  Every base
// component is +-c, and (+-x)*c = +-(x*c) exactly, so the nine products
// Oq[k][j]*c are formed once and each patch component is their signed sum in
// the reference's order (utils.h:17-21).  The patches are built one at a
// time (a scheduling barrier keeps them from being hoisted together), each
// tested against up to two directions: the old (uo, if want_o) and the new
// (un, if want_n) separation of a trial.
__device__ __forceinline__ void kf_any_q2(double uox, double uoy, double uoz, bool want_o, double unx, double uny,
                                          double unz, bool want_n, const double Oq[9], double cosmax, bool &hit_o,
                                          bool &hit_n) {
	double m[9];
#pragma unroll
	for(int e = 0; e < 9; e++) m[e] = Oq[e] * KF_C;
	hit_o = false;
	hit_n = false;
#pragma unroll
	for(int b = 0; b < 4; b++) {
		const bool n0 = (KF_NEG >> (3 * b)) & 1u, n1 = (KF_NEG >> (3 * b + 1)) & 1u, n2 = (KF_NEG >> (3 * b + 2)) & 1u;
		double q[3];
#pragma unroll
		for(int k = 0; k < 3; k++) {
			double a0 = n0 ? -m[3 * k] : m[3 * k];
			double a1 = n1 ? -m[3 * k + 1] : m[3 * k + 1];
			double a2 = n2 ? -m[3 * k + 2] : m[3 * k + 2];
			q[k] = a0 + a1 + a2;
		}
		hit_o |= want_o && (-kf_dot(uox, uoy, uoz, q[0], q[1], q[2]) > cosmax);
		hit_n |= want_n && (-kf_dot(unx, uny, unz, q[0], q[1], q[2]) > cosmax);
		__builtin_amdgcn_sched_barrier(0);
	}
}

__device__ __forceinline__ bool kf_any_q(double ux, double uy, double uz, const double Oq[9], double cosmax) {
	bool h, unused;
	kf_any_q2(ux, uy, uz, true, 0., 0., 0., false, Oq, cosmax, h, unused);
	return h;
}

// ---------------------------------------------------------------------
// Filtered angular tests: the decisions of kf_any_p / kf_any_q without the
// square root and the divisions of d^, and without forming q's patches, for
// all but a vanishing share of pairs.
//
// For a patch vector v (|v| = 1 to rounding) and the minimum-image d with
// r2 = d.d as above, the reference decides (d / sqrt(r2)).v > cosmax; its
// roundings keep that dot within 2e-15 of the exact cosine (d.v)/|d| for
// |d| ~ 1.  Here s ~ d.v is formed in an order of its own: on the p side
// from p's patches by fused multiply-adds (error below 4e-16 |d|), on the q
// side in q's body frame, e = Oq^T d and d.q_b = c (sigma_b . e) for the
// sign pattern sigma_b of patch b (error below 2.2e-15 |d| against the
// reference's rounded patch), so q's patches are never built.  Each side
// takes the largest t = s|s| over the particle's four patches and compares it
// with hi = c2 r2 + eps r2 and lo = c2 r2 - eps r2, c2 = fl(cosmax^2); the
// error of t - c2 r2 against the exact (d.v)|d.v| - cosmax^2 |d|^2 stays below
// 6e-15 r2.  Hence t > hi puts the exact cosine of that patch more than eps/2
// above cosmax (and s > 0), and t < lo puts every patch's as far below (or
// its s <= 0); either way the reference's rounded dots are on the same side
// of cosmax.  Inside the band lo <= t <= hi (one pair in ~10^11) the test
// falls back to the reference's own arithmetic.  eps = 1e-12.  The q side
// scales t by 1/c^2, taken as 3 (exact to 2.3e-16, relative).
// The filter needs cosmax >= KF_FILTER_MIN_COS; for a smaller (or negative)
// cosmax every test takes the exact form.
constexpr double KF_EPS = 1e-12;
constexpr double KF_FILTER_MIN_COS = 1e-3;
constexpr double KF_INV_C2 = 3.;
enum : int { KF_NO = 0, KF_YES = 1, KF_UNSURE = 2 };

// The band of one separation (c2 = fl(cosmax^2), r2 = d.d).
struct KfBand {
	double hi, lo;
};

__device__ __forceinline__ KfBand kf_band(double c2, double r2) {
	const double c = c2 * r2, e = KF_EPS * r2;
	return {c + e, c - e};
}

// The same band in q's units (t / c^2).
__device__ __forceinline__ KfBand kf_band_q(const KfBand &b) { return {KF_INV_C2 * b.hi, KF_INV_C2 * b.lo}; }

// Branch-free classification of the largest t of a side: bit 0 set if
// certainly above, bit 1 set if inside the band (undecided).
__device__ __forceinline__ unsigned kf_fast_class(double t, const KfBand &b) {
	const bool yes = t > b.hi;
	const bool band = !yes && !(t < b.lo);
	return (unsigned) yes | ((unsigned) band << 1);
}

__device__ __forceinline__ double kf_fdot(double dx, double dy, double dz, double a, double b, double c) {
	return __builtin_fma(dx, a, __builtin_fma(dy, b, dz * c));
}

// p side on the raw separation: the largest (d.p_a)|d.p_a| over the patches P
// of p (12 doubles, patch-major), classified
__device__ __forceinline__ unsigned kf_fast_p(double dx, double dy, double dz, const double *P, const KfBand &b) {
	double t = 0.;
#pragma unroll
	for(int a = 0; a < 4; a++) {
		const double s = kf_fdot(dx, dy, dz, P[3 * a], P[3 * a + 1], P[3 * a + 2]);
		const double ta = s * __builtin_fabs(s);
		t = a == 0 ? ta : __builtin_fmax(t, ta);
	}
	return kf_fast_class(t, b);
}

// q side: the largest -(g_b)|g_b| over q's patches, g_b = sigma_b . e with
// e = Oq^T d (so d.q_b = c g_b and the condition -(d.q_b) > cosmax |d| reads
// -(g_b)|g_b| > cosmax^2 r2 / c^2)
__device__ __forceinline__ double kf_q_top(double dx, double dy, double dz, const double Oq[9]) {
	const double e0 = kf_fdot(dx, dy, dz, Oq[0], Oq[3], Oq[6]);
	const double e1 = kf_fdot(dx, dy, dz, Oq[1], Oq[4], Oq[7]);
	const double e2 = kf_fdot(dx, dy, dz, Oq[2], Oq[5], Oq[8]);
	double t = 0.;
#pragma unroll
	for(int b = 0; b < 4; b++) {
		const double a0 = ((KF_NEG >> (3 * b)) & 1u) ? -e0 : e0;
		const double a1 = ((KF_NEG >> (3 * b + 1)) & 1u) ? -e1 : e1;
		const double a2 = ((KF_NEG >> (3 * b + 2)) & 1u) ? -e2 : e2;
		const double g = a0 + a1 + a2;
		const double tb = -(g * __builtin_fabs(g));
		t = b == 0 ? tb : __builtin_fmax(t, tb);
	}
	return t;
}

// q side on one separation (bq = kf_band_q of its band)
__device__ __forceinline__ unsigned kf_fast_q1(double dx, double dy, double dz, const KfBand &bq, const double Oq[9]) {
	return kf_fast_class(kf_q_top(dx, dy, dz, Oq), bq);
}

// q side on the old (o) and the new (n) separation of a trial
__device__ __forceinline__ void kf_fast_q2(double ox, double oy, double oz, const KfBand &boq, double nx, double ny,
                                           double nz, const KfBand &bnq, const double Oq[9], unsigned &fo,
                                           unsigned &fn) {
	fo = kf_fast_q1(ox, oy, oz, boq, Oq);
	fn = kf_fast_q1(nx, ny, nz, bnq, Oq);
}

// Cell coordinate of x on a grid of n cells over [s, s+L) (periodic).
// Deterministic: the rebuild and the in-cell test of a trial move call the
// same function on the same doubles.
__device__ __forceinline__ int cell_coord(double x, double s, double invL, int n) {
	double t = (x - s) * invL;
	t = t - __builtin_floor(t);
	int i = (int) (t * (double) n);
	return i < n ? i : n - 1;
}

// ---------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds, multipliers
// 0xD2511F53 / 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85.
struct u32x4 {
	uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
	for(int i = 0; i < 10; i++) {
		// each 32 x 32 -> 64 product in one v_mad_u64_u32
		const uint64_t p0 = (uint64_t) 0xD2511F53u * c.x, p1 = (uint64_t) 0xCD9E8D57u * c.z;
		const uint32_t hi0 = (uint32_t) (p0 >> 32), lo0 = (uint32_t) p0;
		const uint32_t hi1 = (uint32_t) (p1 >> 32), lo1 = (uint32_t) p1;
		c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	return c;
}

// Uniform double in [0, 1) with 32 random bits.
__device__ __forceinline__ double u01(uint32_t v) { return (double) v * 2.3283064365386962890625e-10; }

// Counter layout (the key is the 64-bit seed):
//   x: cell index within the replica (< 2^24; or 0 for per-replica draws)
//   y: stream tag (8 bits) | colour (3 bits) << 8 | trial index (16 bits) << 16
//   z: sweep index, low 32 bits
//   w: replica (20 bits) | sweep index bits 32..43 << 20
enum : uint32_t {
	TAG_TRIAL_A = 0,
	TAG_TRIAL_B = 1,
	TAG_LEARN_A = 2,
	TAG_LEARN_B = 3,
	TAG_AXIS = 4,  // rejection block index in the high byte of the cell word
	TAG_AXIS_FALLBACK = 5,
	TAG_SHIFT = 8,
	TAG_COLOUR = 9,
	TAG_VOLUME = 10,
	TAG_GC = 11
};

__device__ __forceinline__ u32x4 rng(uint64_t seed, int rep, uint64_t sweep, uint32_t cell, uint32_t tag,
                                     uint32_t colour, uint32_t trial) {
	u32x4 c{cell, tag | (colour << 8) | (trial << 16), (uint32_t) sweep,
	        ((uint32_t) rep & 0xFFFFFu) | ((uint32_t) (sweep >> 32) << 20)};
	return philox(c, (uint32_t) seed, (uint32_t) (seed >> 32));
}

// Random colour order of a sweep: Fisher-Yates over 8 colours, the
// permutation kept as eight 4-bit fields of one word; 16 random bits per
// step from one Philox block.  Returns the colour of pass k.
__host__ __device__ __forceinline__ uint32_t colour_perm_from(const uint32_t w4[4]) {
	uint32_t perm = 0x76543210u;
	for(int i = 7; i > 0; i--) {
		const int s = 7 - i;
		uint32_t w = (w4[s & 3] >> ((s >> 2) * 16)) & 0xFFFFu;
		uint32_t j = (w * (uint32_t) (i + 1)) >> 16;
		uint32_t pi = (perm >> (4 * i)) & 0xFu, pj = (perm >> (4 * j)) & 0xFu;
		perm &= ~((0xFu << (4 * i)) | (0xFu << (4 * j)));
		perm |= (pj << (4 * i)) | (pi << (4 * j));
	}
	return perm;
}

__device__ __forceinline__ int colour_of_pass(uint64_t seed, int rep, uint64_t sweep, int k) {
	u32x4 v = rng(seed, rep, sweep, 0, TAG_COLOUR, 0, 0);
	const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
	return (int) ((colour_perm_from(w4) >> (4 * k)) & 0xFu);
}

// Rodrigues rotation about a unit axis by t (get_rotation_matrix,
// utils.c:186-202).
__device__ __forceinline__ void rotation_matrix(const double a[3], double st, double ct, double R[9]) {
	double uct = 1. - ct;
	R[0] = a[0] * a[0] + (1. - a[0] * a[0]) * ct;
	R[1] = a[0] * a[1] * uct - a[2] * st;
	R[2] = a[0] * a[2] * uct + a[1] * st;
	R[3] = a[0] * a[1] * uct + a[2] * st;
	R[4] = a[1] * a[1] + (1. - a[1] * a[1]) * ct;
	R[5] = a[1] * a[2] * uct - a[0] * st;
	R[6] = a[0] * a[2] * uct - a[1] * st;
	R[7] = a[1] * a[2] * uct + a[0] * st;
	R[8] = a[2] * a[2] + (1. - a[2] * a[2]) * ct;
}

// O' = O R, rows of O (matrix_matrix_multiplication, utils.c:172-184;
// utils_rotate_matrix, utils.c:204-209).
__device__ __forceinline__ void mat_mul3(const double O[9], const double R[9], double On[9]) {
#pragma unroll
	for(int i = 0; i < 3; i++)
#pragma unroll
		for(int j = 0; j < 3; j++) On[3 * i + j] = O[3 * i + 0] * R[j] + O[3 * i + 1] * R[3 + j] + O[3 * i + 2] * R[6 + j];
}

// Gram-Schmidt of the rows, then each normalised (utils.c:91-113, 31-36).
__device__ __forceinline__ void gram_schmidt_rows(double O[9]) {
	double *v1 = O, *v2 = O + 3, *v3 = O + 6;
	double n1 = kf_dot(v1[0], v1[1], v1[2], v1[0], v1[1], v1[2]);
	double c21 = kf_dot(v2[0], v2[1], v2[2], v1[0], v1[1], v1[2]) / n1;
	v2[0] -= c21 * v1[0];
	v2[1] -= c21 * v1[1];
	v2[2] -= c21 * v1[2];
	double s31 = kf_dot(v3[0], v3[1], v3[2], v1[0], v1[1], v1[2]);
	double s32 = kf_dot(v3[0], v3[1], v3[2], v2[0], v2[1], v2[2]);
	double n2 = kf_dot(v2[0], v2[1], v2[2], v2[0], v2[1], v2[2]);
	double c31 = s31 / n1;
	double c32 = s32 / n2;
	v3[0] -= c31 * v1[0] + c32 * v2[0];
	v3[1] -= c31 * v1[1] + c32 * v2[1];
	v3[2] -= c31 * v1[2] + c32 * v2[2];
#pragma unroll
	for(int r = 0; r < 3; r++) {
		double *v = O + 3 * r;
		double len = __builtin_sqrt(kf_dot(v[0], v[1], v[2], v[0], v[1], v[2]));
		v[0] = v[0] / len;
		v[1] = v[1] / len;
		v[2] = v[2] / len;
	}
}

}  // namespace ppd
