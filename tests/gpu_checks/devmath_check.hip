// devmath_check -- the device's double-precision sin, cos, exp, log and sqrt,
// as the sampler calls them, against the host's C library on the ranges the
// sampler feeds them.
//
//   sincos  the trial rotation angle u * Theta_max, u in [0, 1), Theta_max
//           up to pi (MC.c:321-325 draws the angle in [0, Theta_max)), plus
//           small angles down to 1e-12;
//   sqrt    the Marsaglia axis factor 2 sqrt(1 - q), q in [0, 1) (utils.c:
//           115-131), |d| in the pair test (MC.c:244), sqrt(V / Lx) of the
//           box-shape move;
//   exp     exp(-dE / T) of the acceptance test when dE is beyond the
//           replica's table (MC.c:332), exp(ln V') of the volume move and
//           exp(arg) of its acceptance (MC.c:475-499);
//   log     ln V and ln(V' / V) of the volume move (MC.c:474, 497).
//
// The device evaluates each input with the same calls the kernels make; the
// host evaluates it with <cmath> and measures the distance in units in the
// last place (ulp).  Prints one line per function:
//   fn NAME n N exact E maxulp U
// and exits 0 when every U is within the bound given on the command line
// (ulps for sin, cos, exp and log; sqrt must be exact: both sides are
// correctly rounded).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

enum Fn { F_SIN = 0, F_COS, F_SQRT, F_EXP, F_LOG, NFN };
static const char *NAMES[NFN] = {"sin", "cos", "sqrt", "exp", "log"};

__global__ void k_eval(const double *x, int n, int fn, double *y) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) return;
	const double v = x[i];
	double r = 0.;
	if(fn == F_SIN || fn == F_COS) {
		double s, c;
		sincos(v, &s, &c);
		r = fn == F_SIN ? s : c;
	}
	else if(fn == F_SQRT) r = __builtin_sqrt(v);
	else if(fn == F_EXP) r = exp(v);
	else r = log(v);
	y[i] = r;
}

static int64_t ordered(double v) {
	int64_t b;
	std::memcpy(&b, &v, 8);
	return b < 0 ? INT64_MIN - b : b;
}

static uint64_t ulp_distance(double a, double b) {
	if(a == b) return 0;
	if(std::isnan(a) || std::isnan(b)) return UINT64_MAX;
	const int64_t ia = ordered(a), ib = ordered(b);
	return ia > ib ? (uint64_t) ia - (uint64_t) ib : (uint64_t) ib - (uint64_t) ia;
}

static std::vector<double> inputs(int fn, std::mt19937_64 &g) {
	std::uniform_real_distribution<double> u(0., 1.);
	std::vector<double> x;
	const int n = 1 << 20;
	x.reserve(n);
	if(fn == F_SIN || fn == F_COS) {
		for(int i = 0; i < n / 2; i++) x.push_back(u(g) * M_PI);
		for(int i = 0; i < n / 4; i++) x.push_back(u(g) * 0.1);
		for(int i = 0; i < n / 4; i++) x.push_back(std::pow(10., -12. + 11. * u(g)));
	}
	else if(fn == F_SQRT) {
		for(int i = 0; i < n / 2; i++) x.push_back(u(g));
		for(int i = 0; i < n / 4; i++) x.push_back(u(g) * 1.2544);
		for(int i = 0; i < n / 4; i++) x.push_back(std::pow(10., 7. * u(g)));
	}
	else if(fn == F_EXP) {
		for(int i = 0; i < n / 2; i++) x.push_back(-60. * u(g));
		for(int i = 0; i < n / 4; i++) x.push_back(25. * u(g));
		for(int i = 0; i < n / 4; i++) x.push_back(-0.1 + 0.2 * u(g));
	}
	else {
		for(int i = 0; i < n / 2; i++) x.push_back(std::pow(10., 8. * u(g)));
		for(int i = 0; i < n / 2; i++) x.push_back(0.9 + 0.2 * u(g));
	}
	return x;
}

static double host_eval(int fn, double v) {
	switch(fn) {
	case F_SIN: return std::sin(v);
	case F_COS: return std::cos(v);
	case F_SQRT: return std::sqrt(v);
	case F_EXP: return std::exp(v);
	default: return std::log(v);
	}
}

#define CHECK(x)                                                                          \
	do {                                                                                  \
		hipError_t e_ = (x);                                                              \
		if(e_ != hipSuccess) {                                                            \
			std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                  \
			return 2;                                                                     \
		}                                                                                 \
	} while(0)

int main(int argc, char **argv) {
	const uint64_t bound = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 1;
	std::mt19937_64 g(20250611);
	int fail = 0;
	for(int fn = 0; fn < NFN; fn++) {
		const std::vector<double> x = inputs(fn, g);
		const int n = (int) x.size();
		double *dx = nullptr, *dy = nullptr;
		CHECK(hipMalloc(&dx, sizeof(double) * n));
		CHECK(hipMalloc(&dy, sizeof(double) * n));
		CHECK(hipMemcpy(dx, x.data(), sizeof(double) * n, hipMemcpyHostToDevice));
		hipLaunchKernelGGL(k_eval, dim3((n + 255) / 256), dim3(256), 0, 0, dx, n, fn, dy);
		CHECK(hipGetLastError());
		std::vector<double> y(n);
		CHECK(hipMemcpy(y.data(), dy, sizeof(double) * n, hipMemcpyDeviceToHost));
		CHECK(hipFree(dx));
		CHECK(hipFree(dy));
		long long exact = 0;
		uint64_t worst = 0;
		double worst_x = 0.;
		for(int i = 0; i < n; i++) {
			const uint64_t d = ulp_distance(y[i], host_eval(fn, x[i]));
			exact += d == 0;
			if(d > worst) {
				worst = d;
				worst_x = x[i];
			}
		}
		std::printf("fn %s n %d exact %lld maxulp %llu at %.17g\n", NAMES[fn], n, exact, (unsigned long long) worst,
		            worst_x);
		if(worst > (fn == F_SQRT ? 0 : bound)) fail = 1;
	}
	return fail;
}
