// pp_host.cpp -- host-side input producer of libpp_hip.so: the random
// sequential insertion start of the reference's generator (generator.c).
//
// Same algorithm and the same 48-bit LCG stream as the reference program:
// with seed < 0 the stream starts from the LCG's default state (the reference
// generator never seeds), so pp_generate(n, rho, -1) reproduces its
// generated.rrr; seed >= 0 starts it as srand48(seed) would.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/pp.h"

namespace {

struct Lcg48 {
	unsigned short x[3] = {0, 0, 0};
	explicit Lcg48(long seed) {
		if(seed >= 0) {
			unsigned long v = (unsigned long) seed & 0xffffffffUL;
			x[0] = 0x330e;
			x[1] = (unsigned short) (v & 0xffff);
			x[2] = (unsigned short) (v >> 16);
		}
	}
	double next() { return erand48(x); }
};

inline double dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

void unit(double *v) {
	double l = std::sqrt(dot(v, v));
	v[0] /= l;
	v[1] /= l;
	v[2] /= l;
}

// Marsaglia: uniform point on the unit sphere
void sphere(Lcg48 &g, double v[3]) {
	double a, b, q;
	do {
		a = 1. - 2. * g.next();
		b = 1. - 2. * g.next();
		q = a * a + b * b;
	} while(q >= 1.);
	double f = 2. * std::sqrt(1. - q);
	v[0] = a * f;
	v[1] = b * f;
	v[2] = 1. - 2. * q;
}

// rows (z, x, y) of three Gram-Schmidt-orthonormalised sphere points; rows
// (x, z, y) when that frame would be left-handed
void orientation(Lcg48 &g, double O[9]) {
	double x[3], y[3], z[3];
	sphere(g, x);
	sphere(g, y);
	sphere(g, z);
	double n1 = dot(x, x);
	double c = dot(y, x) / n1;
	for(int k = 0; k < 3; k++) y[k] -= c * x[k];
	double s31 = dot(z, x), s32 = dot(z, y), n2 = dot(y, y);
	double c31 = s31 / n1, c32 = s32 / n2;
	for(int k = 0; k < 3; k++) z[k] -= c31 * x[k] + c32 * y[k];
	unit(x);
	unit(y);
	unit(z);
	std::memcpy(O, z, sizeof z);
	std::memcpy(O + 3, x, sizeof x);
	std::memcpy(O + 6, y, sizeof y);
	double det = 0.;
	det += O[0] * O[4] * O[8];
	det += O[1] * O[5] * O[6];
	det += O[2] * O[3] * O[7];
	det -= O[2] * O[4] * O[6];
	det -= O[0] * O[5] * O[7];
	det -= O[1] * O[3] * O[8];
	if(det < 0) {
		std::memcpy(O, x, sizeof x);
		std::memcpy(O + 3, z, sizeof z);
	}
}

}  // namespace

extern "C" pp_status pp_generate(int n, double rho, long seed, double *r, double *O, double box[3]) {
	if(n < 0 || !(rho > 0.) || !r || !O || !box) return PP_ERR_ARG;
	if(rho > 0.7) return PP_ERR_ARG;  // random insertion cannot reach such densities
	const double L = std::pow(n / rho, 1. / 3.);
	box[0] = box[1] = box[2] = L;
	int m = (int) std::floor(L / 1.);
	if(m < 3) m = 3;
	std::vector<int> head((size_t) m * m * m, -1), next(std::max(n, 1), -1);
	auto coord = [&](double x) {
		double t = x / L;
		return (int) ((t - std::floor(t)) * (1. - 2.220446049250313e-16) * m);
	};
	Lcg48 g(seed);
	int done = 0;
	while(done < n) {
		double p[3];
		p[0] = g.next() * L;
		p[1] = g.next() * L;
		p[2] = g.next() * L;
		int cx = coord(p[0]), cy = coord(p[1]), cz = coord(p[2]);
		bool hit = false;
		for(int a = -1; a <= 1 && !hit; a++)
			for(int b = -1; b <= 1 && !hit; b++)
				for(int e = -1; e <= 1 && !hit; e++) {
					int c = (((cx + a + m) % m) * m + (cy + b + m) % m) * m + (cz + e + m) % m;
					for(int q = head[c]; q != -1 && !hit; q = next[q]) {
						double d2 = 0.;
						double d[3];
						for(int k = 0; k < 3; k++) {
							d[k] = r[3 * q + k] - p[k];
							d[k] -= L * std::rint(d[k] / L);
						}
						d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
						hit = d2 < 1.;
					}
				}
		if(hit) continue;
		orientation(g, O + 9 * done);
		std::memcpy(r + 3 * done, p, sizeof p);
		int c = (cx * m + cy) * m + cz;
		next[done] = head[c];
		head[c] = done;
		done++;
	}
	return PP_OK;
}
