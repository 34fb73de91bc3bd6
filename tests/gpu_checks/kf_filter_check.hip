// kf_filter_check -- the colour pass's filtered Kern-Frenkel test against the
// reference arithmetic, on pairs placed at the angular thresholds.
//
// The host builds pairs whose patch cosines sit within a few 1e-12 of
// cosmax (on both sides, on the p and on the q side), plus random pairs in
// the bonding shell, and the device classifies each with the filter
// (bonds_new_sep: certain bond, certain no-bond, or inside the band) and with
// the exact form (bonds_new_exact, the arithmetic of MC.c:242-262).  Every
// pair the filter calls certain must match the exact decision; pairs in the
// band must exist (the band is exercised).  Prints one line:
//   pairs P certain C band B mismatches M
// and exits 0 when M == 0 and B > 0.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "pp_device.hpp"

using namespace ppd;

struct Pair {
	double q[3];   // position of q (p sits at the origin)
	double Op[9];  // orientations
	double Oq[9];
};

__global__ void k_check(const Pair *pairs, int n, double L, double rc2, double cosmax, int *out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) return;
	const Pair &pr = pairs[i];
	Trial t;
	t.px = t.py = t.pz = 0.;
	t.nx = t.ny = t.nz = 0.;
	t.Lx = t.Ly = t.Lz = L;
	t.iLx = t.iLy = t.iLz = 1. / L;
	t.rc2 = rc2;
	t.cosmax = cosmax;
	t.c2 = cosmax * cosmax;
	double P[12];
	{
		double Pm[4][3];
		kf_patches(pr.Op, Pm);
		for(int a = 0; a < 4; a++)
			for(int c = 0; c < 3; c++) P[3 * a + c] = Pm[a][c];
	}
	auto load_o = [&](double *Oq) {
		for(int c = 0; c < 9; c++) Oq[c] = pr.Oq[c];
	};
	double wx, wy, wz;
	const double w2 = sep_new(t, pr.q[0], pr.q[1], pr.q[2], wx, wy, wz);
	int bf = 0;
	bool band = false;
	bonds_new_sep(t, P, wx, wy, wz, w2, load_o, bf, band);
	int be = 0;
	bonds_new_exact(t, P, pr.q[0], pr.q[1], pr.q[2], load_o, be);
	out[i] = band ? 2 : (bf == be ? 0 : 1);
}

// rotation taking unit vector a onto unit vector b (Rodrigues), row-major
static void rot_onto(const double a[3], const double b[3], double R[9]) {
	const double v[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
	const double c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
	const double K[9] = {0, -v[2], v[1], v[2], 0, -v[0], -v[1], v[0], 0};
	double K2[9];
	for(int i = 0; i < 3; i++)
		for(int j = 0; j < 3; j++) K2[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
	const double f = 1. / (1. + c);
	for(int e = 0; e < 9; e++) R[e] = (e % 4 == 0 ? 1. : 0.) + K[e] + K2[e] * f;
}

static void unit(double v[3]) {
	const double s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
	for(int c = 0; c < 3; c++) v[c] /= s;
}

int main() {
	const double delta = 0.12, cosmax = 0.92, L = 10.;
	const double rc2 = (1. + delta) * (1. + delta);
	const int n = 1 << 20;
	std::mt19937_64 g(12345);
	std::uniform_real_distribution<double> U(0., 1.), S(-1., 1.);
	std::normal_distribution<double> N01(0., 1.);
	// the base patches (body frame), as the device forms them
	double base[4][3];
	for(int i = 0; i < 4; i++)
		for(int k = 0; k < 3; k++) base[i][k] = kf_base(i, k);
	std::vector<Pair> pairs(n);
	for(int i = 0; i < n; i++) {
		Pair &p = pairs[i];
		double u[3] = {N01(g), N01(g), N01(g)};
		unit(u);
		const double r = 1. + delta * std::sqrt(U(g)) * 0.999;
		for(int c = 0; c < 3; c++) p.q[c] = r * u[c];
		// patch a of p at angle acos(cosmax + eta) from u, patch b of q from -u
		const int mode = i % 4;  // 0: p at threshold, 1: q at threshold, 2: both, 3: random in the cones
		auto aim = [&](const double axis[3], bool at_threshold, double dir[3]) {
			double cs;
			if(at_threshold) {
				const double eta = std::ldexp(S(g), -(int) (30 + 20 * U(g)));  // |eta| in about [1e-15, 1e-9]
				cs = cosmax + eta;
			}
			else cs = cosmax + (1. - cosmax) * U(g) * 1.2 - 0.1 * (1. - cosmax);
			if(cs > 1.) cs = 1.;
			double w[3] = {N01(g), N01(g), N01(g)};
			const double d = w[0] * axis[0] + w[1] * axis[1] + w[2] * axis[2];
			for(int c = 0; c < 3; c++) w[c] -= d * axis[c];
			unit(w);
			const double sn = std::sqrt(std::max(0., 1. - cs * cs));
			for(int c = 0; c < 3; c++) dir[c] = cs * axis[c] + sn * w[c];
		};
		double dp[3], dq[3], mu[3] = {-u[0], -u[1], -u[2]};
		aim(u, mode == 0 || mode == 2, dp);
		aim(mu, mode == 1 || mode == 2, dq);
		// O maps body base patch a to dp: O = R with R b^_a = dp, then a
		// random spin about dp
		auto orient = [&](const double dir[3], double O[9]) {
			const int a = (int) (U(g) * 4) & 3;
			double b[3] = {base[a][0], base[a][1], base[a][2]};
			unit(b);
			double R[9];
			rot_onto(b, dir, R);
			// spin about dir by a random angle
			const double th = 2. * M_PI * U(g), c = std::cos(th), s = std::sin(th);
			const double K[9] = {0, -dir[2], dir[1], dir[2], 0, -dir[0], -dir[1], dir[0], 0};
			double Sp[9];
			for(int i2 = 0; i2 < 3; i2++)
				for(int j = 0; j < 3; j++) {
					double k2 = K[3 * i2] * K[j] + K[3 * i2 + 1] * K[3 + j] + K[3 * i2 + 2] * K[6 + j];
					Sp[3 * i2 + j] = (i2 == j ? 1. : 0.) + s * K[3 * i2 + j] + (1. - c) * k2;
				}
			// O = Sp R; the device's patch i is O * b_i over the rows of O
			for(int i2 = 0; i2 < 3; i2++)
				for(int j = 0; j < 3; j++) O[3 * i2 + j] = Sp[3 * i2] * R[j] + Sp[3 * i2 + 1] * R[3 + j] + Sp[3 * i2 + 2] * R[6 + j];
		};
		orient(dp, p.Op);
		orient(dq, p.Oq);
	}
	Pair *d_pairs = nullptr;
	int *d_out = nullptr;
	if(hipMalloc(&d_pairs, sizeof(Pair) * n) != hipSuccess || hipMalloc(&d_out, sizeof(int) * n) != hipSuccess) {
		fprintf(stderr, "hipMalloc failed\n");
		return 2;
	}
	if(hipMemcpy(d_pairs, pairs.data(), sizeof(Pair) * n, hipMemcpyHostToDevice) != hipSuccess) return 2;
	hipLaunchKernelGGL(k_check, dim3((n + 255) / 256), dim3(256), 0, 0, d_pairs, n, L, rc2, cosmax, d_out);
	std::vector<int> out(n);
	if(hipMemcpy(out.data(), d_out, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess) {
		fprintf(stderr, "kernel or copy failed\n");
		return 2;
	}
	long certain = 0, band = 0, bad = 0;
	for(int v : out) {
		if(v == 2) band++;
		else {
			certain++;
			bad += v;
		}
	}
	printf("pairs %d certain %ld band %ld mismatches %ld\n", n, certain, band, bad);
	(void) hipFree(d_pairs);
	(void) hipFree(d_out);
	return (bad == 0 && band > 0) ? 0 : 1;
}
