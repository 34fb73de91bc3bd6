// philox_check -- the sampler's own random numbers, evaluated on the device
// by the functions the kernels call (kf_device.hpp), for the test to compare
// with the CPU oracle:
//
//   philox            Philox4x32-10 of a (counter, key) block: the Random123
//                     known-answer inputs, edge words and 2^20 random blocks;
//   rng               the counter layout of a draw: (seed, replica, sweep,
//                     cell, tag, colour, trial) -> counter and key, then
//                     Philox; 2^18 random argument tuples, sweep indices past
//                     2^32 included;
//   colour_perm_from  the colour order of a sweep from one Philox block,
//                     2^18 random blocks.
//
// Writes every input with its device output to the file named on the
// command line, as little-endian uint32 words:
//   n_a, then n_a x (ctr[4], key[2], out[4])
//   n_b, then n_b x (seed lo, seed hi, rep, sweep lo, sweep hi, cell, tag,
//                    colour, trial, out[4])
//   n_c, then n_c x (w[4], perm)
// and prints "philox A rng B perm C" with the three counts.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "pp_device.hpp"

using namespace ppd;

struct RngArgs {
	uint32_t seed_lo, seed_hi, rep, sweep_lo, sweep_hi, cell, tag, colour, trial;
};

__global__ void k_philox(const uint32_t *in, int n, uint32_t *out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) return;
	const uint32_t *c = in + 6 * (size_t) i;
	const u32x4 v = philox(u32x4{c[0], c[1], c[2], c[3]}, c[4], c[5]);
	uint32_t *o = out + 4 * (size_t) i;
	o[0] = v.x;
	o[1] = v.y;
	o[2] = v.z;
	o[3] = v.w;
}

__global__ void k_rng(const RngArgs *in, int n, uint32_t *out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) return;
	const RngArgs a = in[i];
	const uint64_t seed = ((uint64_t) a.seed_hi << 32) | a.seed_lo;
	const uint64_t sweep = ((uint64_t) a.sweep_hi << 32) | a.sweep_lo;
	const u32x4 v = rng(seed, (int) a.rep, sweep, a.cell, a.tag, a.colour, a.trial);
	uint32_t *o = out + 4 * (size_t) i;
	o[0] = v.x;
	o[1] = v.y;
	o[2] = v.z;
	o[3] = v.w;
}

__global__ void k_perm(const uint32_t *in, int n, uint32_t *out) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= n) return;
	const uint32_t w[4] = {in[4 * (size_t) i], in[4 * (size_t) i + 1], in[4 * (size_t) i + 2], in[4 * (size_t) i + 3]};
	out[i] = colour_perm_from(w);
}

template <class Tp>
static bool to_device(const std::vector<Tp> &h, Tp **d) {
	if(hipMalloc(d, sizeof(Tp) * (h.empty() ? 1 : h.size())) != hipSuccess) return false;
	return hipMemcpy(*d, h.data(), sizeof(Tp) * h.size(), hipMemcpyHostToDevice) == hipSuccess;
}

int main(int argc, char **argv) {
	if(argc != 2) {
		fprintf(stderr, "usage: philox_check OUT\n");
		return 2;
	}
	std::mt19937_64 g(20240531);
	auto u32 = [&]() { return (uint32_t) (g() >> 32); };

	// (a) Philox blocks: Random123's known-answer inputs, edge words, random
	std::vector<uint32_t> a_in;
	const uint32_t kat[3][6] = {{0, 0, 0, 0, 0, 0},
	                            {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
	                            {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u}};
	for(const auto &k : kat) a_in.insert(a_in.end(), k, k + 6);
	const uint32_t edge[4] = {0u, 1u, 0x80000000u, 0xffffffffu};
	for(int e = 0; e < 4 * 4; e++)
		for(int b = 0; b < 6; b++) a_in.push_back(b == e % 6 ? edge[e / 4] : edge[(e + b) & 3]);
	const int n_rand = 1 << 20;
	for(int i = 0; i < n_rand; i++)
		for(int b = 0; b < 6; b++) a_in.push_back(u32());
	const int n_a = (int) (a_in.size() / 6);

	// (b) draws: the fields at their full widths (replica 20 bits, tag 8,
	// colour 3, trial 16, cell 24 plus a block index in the high byte,
	// sweep 44 bits)
	const int n_b = 1 << 18;
	std::vector<RngArgs> b_in(n_b);
	for(int i = 0; i < n_b; i++) {
		RngArgs &a = b_in[i];
		a.seed_lo = u32();
		a.seed_hi = (i & 3) == 0 ? 0u : u32();
		a.rep = u32() & 0xFFFFFu;
		a.sweep_lo = u32();
		a.sweep_hi = (i & 1) ? (u32() & 0xFFFu) : 0u;
		a.cell = u32();
		a.tag = u32() & 0xFFu;
		a.colour = u32() & 7u;
		a.trial = u32() & 0xFFFFu;
	}

	// (c) colour orders from random blocks
	const int n_c = 1 << 18;
	std::vector<uint32_t> c_in(4 * (size_t) n_c);
	for(auto &w : c_in) w = u32();

	uint32_t *d_a = nullptr, *d_ao = nullptr, *d_bo = nullptr, *d_c = nullptr, *d_co = nullptr;
	RngArgs *d_b = nullptr;
	if(!to_device(a_in, &d_a) || !to_device(b_in, &d_b) || !to_device(c_in, &d_c) ||
	   hipMalloc(&d_ao, sizeof(uint32_t) * 4 * (size_t) n_a) != hipSuccess ||
	   hipMalloc(&d_bo, sizeof(uint32_t) * 4 * (size_t) n_b) != hipSuccess ||
	   hipMalloc(&d_co, sizeof(uint32_t) * (size_t) n_c) != hipSuccess) {
		fprintf(stderr, "device allocation or upload failed\n");
		return 1;
	}
	hipLaunchKernelGGL(k_philox, dim3((n_a + 255) / 256), dim3(256), 0, 0, d_a, n_a, d_ao);
	hipLaunchKernelGGL(k_rng, dim3((n_b + 255) / 256), dim3(256), 0, 0, d_b, n_b, d_bo);
	hipLaunchKernelGGL(k_perm, dim3((n_c + 255) / 256), dim3(256), 0, 0, d_c, n_c, d_co);
	std::vector<uint32_t> a_out(4 * (size_t) n_a), b_out(4 * (size_t) n_b), c_out(n_c);
	if(hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess ||
	   hipMemcpy(a_out.data(), d_ao, a_out.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
	   hipMemcpy(b_out.data(), d_bo, b_out.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
	   hipMemcpy(c_out.data(), d_co, c_out.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
		fprintf(stderr, "kernel or download failed\n");
		return 1;
	}

	FILE *f = fopen(argv[1], "wb");
	if(!f) {
		fprintf(stderr, "cannot write %s\n", argv[1]);
		return 1;
	}
	bool ok = true;
	auto put = [&](const uint32_t *p, size_t n) { ok = ok && fwrite(p, 4, n, f) == n; };
	const uint32_t hdr_a = (uint32_t) n_a, hdr_b = (uint32_t) n_b, hdr_c = (uint32_t) n_c;
	put(&hdr_a, 1);
	for(int i = 0; i < n_a; i++) {
		put(&a_in[6 * (size_t) i], 6);
		put(&a_out[4 * (size_t) i], 4);
	}
	put(&hdr_b, 1);
	for(int i = 0; i < n_b; i++) {
		put((const uint32_t *) &b_in[i], 9);
		put(&b_out[4 * (size_t) i], 4);
	}
	put(&hdr_c, 1);
	for(int i = 0; i < n_c; i++) {
		put(&c_in[4 * (size_t) i], 4);
		put(&c_out[i], 1);
	}
	ok = (fclose(f) == 0) && ok;
	if(!ok) {
		fprintf(stderr, "writing %s failed\n", argv[1]);
		return 1;
	}
	printf("philox %d rng %d perm %d\n", n_a, n_b, n_c);
	return 0;
}
