// pp_net.cpp -- C ABI of the learned proposal's network: creation (random
// init), weights in and out, and host-driven forward / loss-gradient calls
// that run the device kernels (pp_learn.hip) on host arrays.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pp.h"
#include "pp_net.h"

using namespace ppd;

static_assert(PP_NET_INPUTS == NN_IN && PP_NET_OUTPUTS == NN_OUT && PP_NET_WEIGHTS == NN_W && PP_NET_BIASES == NN_B,
              "the C ABI's network sizes");

pp_status pp_set_error(pp_status s, const char *msg);  // pp_api.cpp

namespace {

pp_status nfail(pp_status s, const char *fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	return pp_set_error(s, buf);
}

#define NCHECK(call)                                                                                        \
	do {                                                                                                    \
		hipError_t e_ = (call);                                                                             \
		if(e_ != hipSuccess) return nfail(PP_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));       \
	} while(0)

template <typename Tp>
pp_status nalloc(pp_net *n, Tp **p, size_t count) {
	void *q = nullptr;
	if(hipMalloc(&q, count * sizeof(Tp)) != hipSuccess) return nfail(PP_ERR_NOMEM, "hipMalloc(%zu bytes)", count * sizeof(Tp));
	n->allocs.push_back(q);
	*p = (Tp *) q;
	return PP_OK;
}

// round to nearest even
uint16_t to_bf16(float f) {
	uint32_t u;
	std::memcpy(&u, &f, 4);
	if((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t) ((u >> 16) | 0x40);  // NaN stays NaN
	u += 0x7fffu + ((u >> 16) & 1u);
	return (uint16_t) (u >> 16);
}

float from_bf16(uint16_t b) {
	uint32_t u = (uint32_t) b << 16;
	float f;
	std::memcpy(&f, &u, 4);
	return f;
}

// SplitMix64 (Steele, Lea, Flood 2014) for the host-side initial weights
struct Mix64 {
	uint64_t s;
	uint64_t next() {
		uint64_t z = (s += 0x9E3779B97F4A7C15ull);
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	}
	double uniform() { return (double) (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// the real (unpadded) fan-in / fan-out of layer l
int fan_in(int l) { return l == 0 ? NN_IN : NN_H; }
int fan_out(int l) { return l == NN_LAYERS - 1 ? NN_OUT : NN_H; }

pp_status push(pp_net *n, const std::vector<float> &w, const std::vector<float> &b) {
	NCHECK(hipMemcpy(n->d.w, w.data(), sizeof(float) * NN_W, hipMemcpyHostToDevice));
	NCHECK(hipMemcpy(n->d.b, b.data(), sizeof(float) * NN_B, hipMemcpyHostToDevice));
	launch_net_pack(n->d, n->stream);
	NCHECK(hipStreamSynchronize(n->stream));
	return PP_OK;
}

}  // namespace

pp_net::~pp_net() {
	for(void *p : allocs) (void) hipFree(p);
	if(stream) (void) hipStreamDestroy(stream);
}

int ppd::net_reserve(pp_net *n, int m) {
	const int need = (m + 31) & ~31;
	if(need <= n->cap) return PP_OK;
	// the smaller scratch of earlier calls goes (the stream is idle: every
	// call that used it synchronised)
	void *old[] = {n->x, n->h, n->dy, n->dh, n->y, n->target, n->loss};
	for(void *p : old) {
		if(!p) continue;
		for(size_t k = 0; k < n->allocs.size(); k++)
			if(n->allocs[k] == p) {
				(void) hipFree(p);
				n->allocs.erase(n->allocs.begin() + (long) k);
				break;
			}
	}
	n->x = n->h = n->dy = n->dh = nullptr;
	n->y = n->target = n->loss = nullptr;
	n->cap = 0;
	const size_t r = (size_t) need;
	pp_status st;
	if((st = nalloc(n, &n->x, r * NN_IN_PAD)) || (st = nalloc(n, &n->h, 3 * r * NN_H)) ||
	   (st = nalloc(n, &n->dy, r * NN_OUT_PAD)) || (st = nalloc(n, &n->dh, 3 * r * NN_H)) ||
	   (st = nalloc(n, &n->y, r * NN_OUT_PAD)) || (st = nalloc(n, &n->target, r * 6)) || (st = nalloc(n, &n->loss, r)))
		return st;
	n->cap = need;
	return PP_OK;
}

extern "C" {

pp_status pp_net_create(int device, uint64_t seed, pp_net **out) {
	if(!out) return nfail(PP_ERR_ARG, "out is NULL");
	*out = nullptr;
	int ndev = 0;
	if(hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return nfail(PP_ERR_HIP, "no HIP device available");
	if(device < 0 || device >= ndev) return nfail(PP_ERR_ARG, "device %d out of range", device);
	pp_net *n = new pp_net();
	n->device = device;
	auto bail = [&](pp_status s) {
		delete n;
		return s;
	};
	if(hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
		return bail(nfail(PP_ERR_HIP, "stream creation failed"));
	NetDev &d = n->d;
	pp_status st;
	if((st = nalloc(n, &d.w, NN_W)) || (st = nalloc(n, &d.b, NN_B)) || (st = nalloc(n, &d.mw, NN_W)) ||
	   (st = nalloc(n, &d.vw, NN_W)) || (st = nalloc(n, &d.mb, NN_B)) || (st = nalloc(n, &d.vb, NN_B)) ||
	   (st = nalloc(n, &d.gw, NN_W)) || (st = nalloc(n, &d.gb, NN_B)) || (st = nalloc(n, &d.frag, NN_W)) ||
	   (st = nalloc(n, &d.fragT, NN_W)))
		return bail(st);
	// Glorot-uniform weights on the real (unpadded) entries; the last layer
	// scaled by 0.1 and its biases set so that the untrained proposal is close
	// to the plain move: displacement sigma 0.03, almost uniform axis
	Mix64 g{seed};
	std::vector<float> w(NN_W, 0.f), b(NN_B, 0.f);
	for(int l = 0; l < NN_LAYERS; l++) {
		const double lim = std::sqrt(6.0 / (fan_in(l) + fan_out(l))) * (l == NN_LAYERS - 1 ? 0.1 : 1.0);
		for(int k = 0; k < fan_in(l); k++)
			for(int j = 0; j < fan_out(l); j++)
				w[nn_woff(l) + k * nn_n(l) + j] = (float) ((2. * g.uniform() - 1.) * lim);
	}
	float *bl = b.data() + nn_boff(NN_LAYERS - 1);
	for(int k = 0; k < 3; k++) bl[3 + k] = std::log(0.03f);
	for(int c = 0; c < NN_MIX; c++) {
		for(int k = 0; k < 3; k++) bl[6 + 5 * c + k] = (float) (2. * g.uniform() - 1.);
		bl[6 + 5 * c + 3] = std::log(0.1f);
	}
	std::vector<float> zw(NN_W, 0.f), zb(NN_B, 0.f);
	if(hipMemcpy(d.mw, zw.data(), sizeof(float) * NN_W, hipMemcpyHostToDevice) != hipSuccess ||
	   hipMemcpy(d.vw, zw.data(), sizeof(float) * NN_W, hipMemcpyHostToDevice) != hipSuccess ||
	   hipMemcpy(d.mb, zb.data(), sizeof(float) * NN_B, hipMemcpyHostToDevice) != hipSuccess ||
	   hipMemcpy(d.vb, zb.data(), sizeof(float) * NN_B, hipMemcpyHostToDevice) != hipSuccess)
		return bail(nfail(PP_ERR_HIP, "initial upload failed"));
	if((st = push(n, w, b))) return bail(st);
	*out = n;
	return PP_OK;
}

pp_status pp_net_set(pp_net *n, const float *w, const float *b) {
	if(!n || !w || !b) return nfail(PP_ERR_ARG, "NULL argument");
	NCHECK(hipSetDevice(n->device));
	std::vector<float> hw(w, w + NN_W), hb(b, b + NN_B);
	return push(n, hw, hb);
}

pp_status pp_net_get(pp_net *n, float *w, float *b) {
	if(!n) return nfail(PP_ERR_ARG, "net is NULL");
	NCHECK(hipSetDevice(n->device));
	NCHECK(hipStreamSynchronize(n->stream));
	if(w) NCHECK(hipMemcpy(w, n->d.w, sizeof(float) * NN_W, hipMemcpyDeviceToHost));
	if(b) NCHECK(hipMemcpy(b, n->d.b, sizeof(float) * NN_B, hipMemcpyDeviceToHost));
	return PP_OK;
}

static pp_status upload_inputs(pp_net *n, int m, const float *x) {
	const int m_pad = (m + 31) & ~31;
	std::vector<uint16_t> hx((size_t) m_pad * NN_IN_PAD, 0);
	for(int i = 0; i < m; i++)
		for(int e = 0; e < NN_IN; e++) hx[(size_t) i * NN_IN_PAD + e] = to_bf16(x[(size_t) i * NN_IN + e]);
	NCHECK(hipMemcpy(n->x, hx.data(), hx.size() * 2, hipMemcpyHostToDevice));
	return PP_OK;
}

pp_status pp_net_forward(pp_net *n, int m, const float *x, float *y) {
	if(!n || m < 0 || (m > 0 && (!x || !y))) return nfail(PP_ERR_ARG, "bad arguments");
	if(m == 0) return PP_OK;
	NCHECK(hipSetDevice(n->device));
	pp_status st;
	if((st = (pp_status) net_reserve(n, m)) || (st = upload_inputs(n, m, x))) return st;
	launch_net_forward(n->d, n->x, m, nullptr, n->y, nullptr, n->stream);
	NCHECK(hipStreamSynchronize(n->stream));
	NCHECK(hipGetLastError());
	std::vector<float> hy((size_t) m * NN_OUT_PAD);
	NCHECK(hipMemcpy(hy.data(), n->y, hy.size() * 4, hipMemcpyDeviceToHost));
	for(int i = 0; i < m; i++)
		for(int e = 0; e < NN_OUT; e++) y[(size_t) i * NN_OUT + e] = hy[(size_t) i * NN_OUT_PAD + e];
	return PP_OK;
}

pp_status pp_net_loss_grad(pp_net *n, int m, const float *x, const float *target, double *loss, float *gw, float *gb) {
	if(!n || m <= 0 || !x || !target) return nfail(PP_ERR_ARG, "bad arguments");
	NCHECK(hipSetDevice(n->device));
	pp_status st;
	if((st = (pp_status) net_reserve(n, m)) || (st = upload_inputs(n, m, x))) return st;
	NCHECK(hipMemcpy(n->target, target, sizeof(float) * 6 * (size_t) m, hipMemcpyHostToDevice));
	launch_net_loss_grad(n->d, n->x, n->target, m, n->h, n->y, n->dy, n->dh, n->loss, n->stream);
	NCHECK(hipStreamSynchronize(n->stream));
	NCHECK(hipGetLastError());
	if(loss) {
		std::vector<float> hl(m);
		NCHECK(hipMemcpy(hl.data(), n->loss, sizeof(float) * m, hipMemcpyDeviceToHost));
		double s = 0.;
		for(float v : hl) s += v;
		*loss = s / m;
	}
	if(gw) NCHECK(hipMemcpy(gw, n->d.gw, sizeof(float) * NN_W, hipMemcpyDeviceToHost));
	if(gb) NCHECK(hipMemcpy(gb, n->d.gb, sizeof(float) * NN_B, hipMemcpyDeviceToHost));
	return PP_OK;
}

pp_status pp_net_descriptor(int device, int n, const double *r, const double *O, const double box[3], int m,
                            const int *idx, float *x) {
	if(n < 1 || m < 0 || !r || !O || !box || (m > 0 && (!idx || !x))) return nfail(PP_ERR_ARG, "bad arguments");
	for(int t = 0; t < m; t++)
		if(idx[t] < 0 || idx[t] >= n) return nfail(PP_ERR_ARG, "idx[%d] = %d out of range", t, idx[t]);
	for(int k = 0; k < 3; k++)
		if(!(box[k] > 0.)) return nfail(PP_ERR_ARG, "box must be positive");
	if(m == 0) return PP_OK;
	NCHECK(hipSetDevice(device));
	Grid g{};
	set_grid_box(g, box);
	double *dr = nullptr, *dO = nullptr;
	int *di = nullptr;
	__bf16 *dx = nullptr;
	pp_status st = PP_OK;
	if(hipMalloc(&dr, sizeof(double) * 3 * n) != hipSuccess || hipMalloc(&dO, sizeof(double) * 9 * n) != hipSuccess ||
	   hipMalloc(&di, sizeof(int) * m) != hipSuccess || hipMalloc(&dx, sizeof(__bf16) * NN_IN_PAD * (size_t) m) != hipSuccess)
		st = nfail(PP_ERR_NOMEM, "hipMalloc failed");
	std::vector<uint16_t> hx((size_t) m * NN_IN_PAD);
	if(!st && (hipMemcpy(dr, r, sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess ||
	           hipMemcpy(dO, O, sizeof(double) * 9 * n, hipMemcpyHostToDevice) != hipSuccess ||
	           hipMemcpy(di, idx, sizeof(int) * m, hipMemcpyHostToDevice) != hipSuccess))
		st = nfail(PP_ERR_HIP, "upload failed");
	if(!st) {
		launch_descriptor(n, dr, dO, g, m, di, dx, nullptr);
		if(hipDeviceSynchronize() != hipSuccess || hipGetLastError() != hipSuccess ||
		   hipMemcpy(hx.data(), dx, hx.size() * 2, hipMemcpyDeviceToHost) != hipSuccess)
			st = nfail(PP_ERR_HIP, "descriptor kernel failed");
	}
	(void) hipFree(dr);
	(void) hipFree(dO);
	(void) hipFree(di);
	(void) hipFree(dx);
	if(st) return st;
	for(int t = 0; t < m; t++)
		for(int e = 0; e < NN_IN; e++) x[(size_t) t * NN_IN + e] = from_bf16(hx[(size_t) t * NN_IN_PAD + e]);
	return PP_OK;
}

void pp_net_destroy(pp_net *n) {
	if(!n) return;
	(void) hipSetDevice(n->device);
	if(n->stream) (void) hipStreamSynchronize(n->stream);
	delete n;
}

}  // extern "C"
