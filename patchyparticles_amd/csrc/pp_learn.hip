// pp_learn.hip -- the learned proposal of the rototranslation sweep (gfx950).
//
// A move of particle p becomes: descriptor of p's neighbourhood -> MLP (BF16
// weights, FP32 accumulation on MFMA) -> a Gaussian displacement in p's body
// frame and a 4-component von Mises-Fisher mixture over the rotation axis (the
// angle is uniform in [0, theta_max), as the reference's) -> the proposed
// state -> the descriptor there -> the MLP again for the reverse density ->
// Metropolis-Hastings with the true Kern-Frenkel energy change and the
// forward/reverse proposal densities:
//   accept with min(1, exp(-dE/T) q(x'->x) / q(x->x')),
// so the stationary distribution is the reference's.  The rotation density
// factor of the angle is the same both ways (reverse = angle theta about -a in
// the body frame of O'), so only the axis densities enter.
//
// The same network is trained here: denoising samples from the sampler's
// states, analytic gradient of -log q, backward passes on MFMA, Adam.
#pragma clang fp contract(off)

#include "pp_device.hpp"

namespace ppd {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr float LOG_2PI = 1.8378770664093453f;

// ------------------------------------------------------------------------
// proposal density: decode the network output, log q, gradient of -log q

struct Prop {
	float mu[3], ls[3];
	float dir[NN_MIX][3], nrm[NN_MIX], lk[NN_MIX], lw[NN_MIX];
	bool ls_in[3], lk_in[NN_MIX];
};

__device__ inline void prop_decode(const float *y, Prop &p) {
	for(int k = 0; k < 3; k++) {
		p.mu[k] = y[k];
		const float v = y[3 + k];
		p.ls_in[k] = v > NN_LS_MIN && v < NN_LS_MAX;
		p.ls[k] = fminf(fmaxf(v, NN_LS_MIN), NN_LS_MAX);
	}
	float mx = -3.0e38f;
	for(int c = 0; c < NN_MIX; c++) mx = fmaxf(mx, y[6 + 5 * c + 4]);
	float se = 0.f;
	for(int c = 0; c < NN_MIX; c++) se += expf(y[6 + 5 * c + 4] - mx);
	const float lse = mx + logf(se);
	for(int c = 0; c < NN_MIX; c++) {
		const float *u = y + 6 + 5 * c;
		const float n = sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + 1e-12f);
		for(int j = 0; j < 3; j++) p.dir[c][j] = u[j] / n;
		p.nrm[c] = n;
		const float v = u[3];
		p.lk_in[c] = v > NN_LK_MIN && v < NN_LK_MAX;
		p.lk[c] = fminf(fmaxf(v, NN_LK_MIN), NN_LK_MAX);
		p.lw[c] = u[4] - lse;
	}
}

// log density of the vMF distribution on S^2 at cosine c to the mean:
// kappa / (4 pi sinh kappa) exp(kappa c)
__device__ inline float log_vmf(float lk, float c) {
	const float k = expf(lk);
	return lk - LOG_2PI - k - logf(-expm1f(-2.f * k)) + k * c;
}

__device__ inline float prop_logq(const Prop &p, const float d[3], const float a[3]) {
	float lg = 0.f;
	for(int k = 0; k < 3; k++) {
		const float z = (d[k] - p.mu[k]) * expf(-p.ls[k]);
		lg += -0.5f * z * z - p.ls[k] - 0.5f * LOG_2PI;
	}
	float l[NN_MIX], mx = -3.0e38f;
	for(int c = 0; c < NN_MIX; c++) {
		l[c] = p.lw[c] + log_vmf(p.lk[c], p.dir[c][0] * a[0] + p.dir[c][1] * a[1] + p.dir[c][2] * a[2]);
		mx = fmaxf(mx, l[c]);
	}
	float se = 0.f;
	for(int c = 0; c < NN_MIX; c++) se += expf(l[c] - mx);
	return lg + mx + logf(se);
}

// -log q and its gradient with respect to the raw outputs y (g[NN_OUT_PAD])
__device__ inline float prop_loss_grad(const float *y, const float d[3], const float a[3], float *g) {
	Prop p;
	prop_decode(y, p);
	float loss = 0.f;
	for(int k = 0; k < 3; k++) {
		const float s2inv = expf(-2.f * p.ls[k]);
		const float diff = d[k] - p.mu[k];
		const float z2 = diff * diff * s2inv;
		loss += 0.5f * z2 + p.ls[k] + 0.5f * LOG_2PI;
		g[k] = -diff * s2inv;
		g[3 + k] = p.ls_in[k] ? 1.f - z2 : 0.f;
	}
	float l[NN_MIX], cs[NN_MIX], mx = -3.0e38f;
	for(int c = 0; c < NN_MIX; c++) {
		cs[c] = p.dir[c][0] * a[0] + p.dir[c][1] * a[1] + p.dir[c][2] * a[2];
		l[c] = p.lw[c] + log_vmf(p.lk[c], cs[c]);
		mx = fmaxf(mx, l[c]);
	}
	float se = 0.f;
	for(int c = 0; c < NN_MIX; c++) se += expf(l[c] - mx);
	const float lse = mx + logf(se);
	loss -= lse;
	for(int c = 0; c < NN_MIX; c++) {
		const float r = expf(l[c] - lse);  // responsibility
		const float w = expf(p.lw[c]);     // mixture weight
		const float k = expf(p.lk[c]);
		const float kcoth = k * (1.f + expf(-2.f * k)) / (-expm1f(-2.f * k));
		float *gc = g + 6 + 5 * c;
		for(int j = 0; j < 3; j++) gc[j] = -r * k * (a[j] - cs[c] * p.dir[c][j]) / p.nrm[c];
		gc[3] = p.lk_in[c] ? -r * (1.f + k * cs[c] - kcoth) : 0.f;
		gc[4] = w - r;
	}
	for(int e = NN_OUT; e < NN_OUT_PAD; e++) g[e] = 0.f;
	return loss;
}

// ------------------------------------------------------------------------
// descriptor

// Quaternion (x, y, z, w), w >= 0, of the rotation matrix M (rows).
__device__ inline void quat_of(const double M[9], float q[4]) {
	double x, y, z, w;
	const double tr = M[0] + M[4] + M[8];
	if(tr > 0.) {
		const double S = sqrt(tr + 1.) * 2.;
		w = 0.25 * S;
		x = (M[7] - M[5]) / S;
		y = (M[2] - M[6]) / S;
		z = (M[3] - M[1]) / S;
	}
	else if(M[0] > M[4] && M[0] > M[8]) {
		const double S = sqrt(1. + M[0] - M[4] - M[8]) * 2.;
		w = (M[7] - M[5]) / S;
		x = 0.25 * S;
		y = (M[1] + M[3]) / S;
		z = (M[2] + M[6]) / S;
	}
	else if(M[4] > M[8]) {
		const double S = sqrt(1. + M[4] - M[0] - M[8]) * 2.;
		w = (M[2] - M[6]) / S;
		x = (M[1] + M[3]) / S;
		y = 0.25 * S;
		z = (M[5] + M[7]) / S;
	}
	else {
		const double S = sqrt(1. + M[8] - M[0] - M[4]) * 2.;
		w = (M[3] - M[1]) / S;
		x = (M[2] + M[6]) / S;
		y = (M[5] + M[7]) / S;
		z = 0.25 * S;
	}
	const double sg = w < 0. ? -1. : 1.;
	q[0] = (float) (sg * x);
	q[1] = (float) (sg * y);
	q[2] = (float) (sg * z);
	q[3] = (float) (sg * w);
}

// Descriptor of a particle at (p, O) among candidate neighbours (n of them;
// pos(i, q[3]) and ori(i, Oq[9]) give candidate i): the NN_NB nearest within
// NN_RCUT (minimum image), nearest first, each as its separation in the body
// frame of O and the quaternion of Oq O^T; zero-padded to NN_IN_PAD.
template <class Pos, class Ori>
__device__ void descriptor(const double p[3], const double O[9], int n, Pos pos, Ori ori, const double L[3],
                           const double iL[3], __bf16 *out) {
	float br2[NN_NB];
	int bi[NN_NB];
	int nb = 0;
	const double rc2 = NN_RCUT * NN_RCUT;
	for(int i = 0; i < n; i++) {
		double q[3];
		pos(i, q);
		double d[3];
		for(int k = 0; k < 3; k++) d[k] = kf_min_image(q[k] - p[k], L[k], iL[k]);
		const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
		if(!(r2 < rc2)) continue;
		const float f = (float) r2;
		if(nb == NN_NB && f >= br2[NN_NB - 1]) continue;
		int at = nb < NN_NB ? nb : NN_NB - 1;
		while(at > 0 && br2[at - 1] > f) {
			br2[at] = br2[at - 1];
			bi[at] = bi[at - 1];
			at--;
		}
		br2[at] = f;
		bi[at] = i;
		if(nb < NN_NB) nb++;
	}
	for(int t = 0; t < NN_NB; t++) {
		float f[NN_FEAT] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
		if(t < nb) {
			double q[3], Oq[9];
			pos(bi[t], q);
			ori(bi[t], Oq);
			double d[3];
			for(int k = 0; k < 3; k++) d[k] = kf_min_image(q[k] - p[k], L[k], iL[k]);
			for(int k = 0; k < 3; k++) f[k] = (float) (O[3 * k] * d[0] + O[3 * k + 1] * d[1] + O[3 * k + 2] * d[2]);
			double M[9];
			for(int i = 0; i < 3; i++)
				for(int j = 0; j < 3; j++) M[3 * i + j] = Oq[3 * i] * O[3 * j] + Oq[3 * i + 1] * O[3 * j + 1] + Oq[3 * i + 2] * O[3 * j + 2];
			quat_of(M, f + 3);
		}
		for(int e = 0; e < NN_FEAT; e++) out[t * NN_FEAT + e] = (__bf16) f[e];
	}
	for(int e = NN_IN; e < NN_IN_PAD; e++) out[e] = (__bf16) 0.f;
}

// ------------------------------------------------------------------------
// the MLP on MFMA (v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][(l/16)*8+i],
// B[(l/16)*8+i][l%16], i < 8, and C[(l/16)*4+j][l%16], j < 4)

__host__ __device__ inline size_t frag_at(int l, int kb, int nt, int lane) {
	return (size_t) nn_woff(l) + ((size_t) (kb * (nn_n(l) / 16) + nt) * 64 + lane) * 8;
}
// fragments of W^T (K' = nn_n(l) rows, N' = nn_k(l) columns)
__host__ __device__ inline size_t fragT_at(int l, int kb, int nt, int lane) {
	return (size_t) nn_woff(l) + ((size_t) (kb * (nn_k(l) / 16) + nt) * 64 + lane) * 8;
}

constexpr int ACT_LD = NN_H + 8;  // LDS row stride of the activations (bf16)

// One wave per 16 rows: every layer of the MLP, activations in LDS.  m rows
// (or *m_dev if given); y: [rows][NN_OUT_PAD]; h (optional): the three
// hidden layers, [3][hs][NN_H], for the backward pass.
__global__ __launch_bounds__(64) void k_net_fwd(NetDev net, const __bf16 *x, int m, const int *m_dev, float *y,
                                                __bf16 *h, int hs) {
	__shared__ __attribute__((aligned(16))) __bf16 act[2][16 * ACT_LD];
	const int lane = threadIdx.x;
	const int rows = m_dev ? *m_dev : m;
	const int row0 = blockIdx.x * 16;
	if(row0 >= rows) return;
	const int r = lane & 15, g = lane >> 4;
	bf16x8 a0[NN_IN_PAD / 32];
	for(int kb = 0; kb < NN_IN_PAD / 32; kb++) {
		bf16x8 v = {};
		if(row0 + r < rows) v = *(const bf16x8 *) (x + (size_t) (row0 + r) * NN_IN_PAD + kb * 32 + g * 8);
		a0[kb] = v;
	}
	int cur = 0;
	for(int l = 0; l < NN_LAYERS; l++) {
		const int nkb = nn_k(l) / 32, nnt = nn_n(l) / 16;
		for(int nt = 0; nt < nnt; nt++) {
			f32x4 acc = {0.f, 0.f, 0.f, 0.f};
			for(int kb = 0; kb < nkb; kb++) {
				const bf16x8 a = l == 0 ? a0[kb] : *(const bf16x8 *) (&act[cur][r * ACT_LD + kb * 32 + g * 8]);
				const bf16x8 b = *(const bf16x8 *) (net.frag + frag_at(l, kb, nt, lane));
				acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
			}
			const int col = nt * 16 + r;
			const float bias = net.b[nn_boff(l) + col];
			for(int j = 0; j < 4; j++) {
				const int row = g * 4 + j;
				float v = acc[j] + bias;
				if(l < NN_LAYERS - 1) {
					v = fmaxf(v, 0.f);
					const __bf16 bv = (__bf16) v;
					act[cur ^ 1][row * ACT_LD + col] = bv;
					if(h && row0 + row < rows) h[((size_t) l * hs + row0 + row) * NN_H + col] = bv;
				}
				else if(row0 + row < rows)
					y[(size_t) (row0 + row) * NN_OUT_PAD + col] = v;
			}
		}
		__syncthreads();
		cur ^= 1;
	}
}

// -log q of each training row (and its gradient, scaled by 1/m, as BF16 for
// the backward kernels); rows in [m, m_pad) get zero gradients.
__global__ void k_net_loss(const float *y, const float *target, int m, int m_pad, __bf16 *dy, float *loss) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= m_pad) return;
	float g[NN_OUT_PAD];
	float l = 0.f;
	if(i < m) {
		const float *t = target + (size_t) i * 6;
		l = prop_loss_grad(y + (size_t) i * NN_OUT_PAD, t, t + 3, g);
		for(int e = 0; e < NN_OUT_PAD; e++) g[e] *= 1.f / (float) m;
	}
	else
		for(int e = 0; e < NN_OUT_PAD; e++) g[e] = 0.f;
	for(int e = 0; e < NN_OUT_PAD; e++) dy[(size_t) i * NN_OUT_PAD + e] = (__bf16) g[e];
	loss[i] = l;
}

// One wave per 16 rows: dH3 = (dY W4^T) * [H3 > 0], dH2 = (dH3 W3^T) * [H2 > 0],
// dH1 = (dH2 W2^T) * [H1 > 0]; dh: [3][hs][NN_H].
__global__ __launch_bounds__(64) void k_net_bwd(NetDev net, const __bf16 *dy, const __bf16 *h, int m_pad, int hs,
                                                __bf16 *dh) {
	__shared__ __attribute__((aligned(16))) __bf16 act[2][16 * ACT_LD];
	const int lane = threadIdx.x;
	const int row0 = blockIdx.x * 16;
	if(row0 >= m_pad) return;
	const int r = lane & 15, g = lane >> 4;
	const bf16x8 gy = *(const bf16x8 *) (dy + (size_t) (row0 + r) * NN_OUT_PAD + g * 8);
	int cur = 0;
	for(int l = NN_LAYERS - 1; l >= 1; l--) {
		// input gradient of layer l: K' = nn_n(l), N' = nn_k(l) = NN_H
		const int nkb = nn_n(l) / 32, nnt = nn_k(l) / 16;
		for(int nt = 0; nt < nnt; nt++) {
			f32x4 acc = {0.f, 0.f, 0.f, 0.f};
			for(int kb = 0; kb < nkb; kb++) {
				const bf16x8 a = l == NN_LAYERS - 1 ? gy : *(const bf16x8 *) (&act[cur][r * ACT_LD + kb * 32 + g * 8]);
				const bf16x8 b = *(const bf16x8 *) (net.fragT + fragT_at(l, kb, nt, lane));
				acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
			}
			const int col = nt * 16 + r;
			for(int j = 0; j < 4; j++) {
				const int row = g * 4 + j;
				const size_t at = ((size_t) (l - 1) * hs + row0 + row) * NN_H + col;
				const float mask = (float) h[at] > 0.f ? 1.f : 0.f;
				const __bf16 v = (__bf16) (acc[j] * mask);
				act[cur ^ 1][row * ACT_LD + col] = v;
				dh[at] = v;
			}
		}
		__syncthreads();
		cur ^= 1;
	}
}

// Weight gradients dW_l = A_l^T dZ_l summed over the m_pad rows (A: x, H1,
// H2, H3; dZ: dH1, dH2, dH3, dY); one wave per 16 x 16 tile of one layer.
__global__ __launch_bounds__(64) void k_net_wgrad(NetDev net, const __bf16 *x, const __bf16 *h, const __bf16 *dy,
                                                  const __bf16 *dh, int m_pad, int hs) {
	int t = blockIdx.x, l = 0;
	while(l < NN_LAYERS && t >= (nn_k(l) / 16) * (nn_n(l) / 16)) {
		t -= (nn_k(l) / 16) * (nn_n(l) / 16);
		l++;
	}
	if(l >= NN_LAYERS) return;
	const int nnt = nn_n(l) / 16;
	const int kt = t / nnt, nt = t % nnt;
	const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
	const __bf16 *A = l == 0 ? x : h + (size_t) (l - 1) * hs * NN_H;
	const int lda = l == 0 ? NN_IN_PAD : NN_H;
	const __bf16 *Z = l == NN_LAYERS - 1 ? dy : dh + (size_t) l * hs * NN_H;
	const int ldz = l == NN_LAYERS - 1 ? NN_OUT_PAD : NN_H;
	f32x4 acc = {0.f, 0.f, 0.f, 0.f};
	for(int c = 0; c < m_pad; c += 32) {
		bf16x8 a, b;
		for(int i = 0; i < 8; i++) {
			const size_t row = (size_t) (c + g * 8 + i);
			a[i] = A[row * lda + kt * 16 + r];
			b[i] = Z[row * ldz + nt * 16 + r];
		}
		acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
	}
	for(int j = 0; j < 4; j++) net.gw[nn_woff(l) + (size_t) (kt * 16 + g * 4 + j) * nn_n(l) + nt * 16 + r] = acc[j];
}

__global__ void k_net_bgrad(NetDev net, const __bf16 *dy, const __bf16 *dh, int m_pad, int hs) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if(i >= NN_B) return;
	int l = 0;
	while(l + 1 < NN_LAYERS && i >= nn_boff(l + 1)) l++;
	const int n = i - nn_boff(l);
	const __bf16 *Z = l == NN_LAYERS - 1 ? dy : dh + (size_t) l * hs * NN_H;
	const int ldz = l == NN_LAYERS - 1 ? NN_OUT_PAD : NN_H;
	float s = 0.f;
	for(int row = 0; row < m_pad; row++) s += (float) Z[(size_t) row * ldz + n];
	net.gb[i] = s;
}

// BF16 operand fragments of W (forward) and W^T (backward) from the FP32 weights.
__global__ void k_net_pack(NetDev net) {
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if(e >= NN_W) return;
	int l = 0;
	while(l + 1 < NN_LAYERS && e >= nn_woff(l + 1)) l++;
	const int loc = e - nn_woff(l), N = nn_n(l);
	const int k = loc / N, n = loc % N;
	const __bf16 v = (__bf16) net.w[e];
	net.frag[frag_at(l, k / 32, n / 16, ((k % 32) / 8) * 16 + n % 16) + k % 8] = v;
	net.fragT[fragT_at(l, n / 32, k / 16, ((n % 32) / 8) * 16 + k % 16) + n % 8] = v;
}

// Adam (Kingma & Ba) on every weight and bias; step t >= 1.
__global__ void k_net_adam(NetDev net, int t, float lr) {
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if(e >= NN_W + NN_B) return;
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	const bool isw = e < NN_W;
	const int i = isw ? e : e - NN_W;
	float *w = isw ? net.w : net.b, *m = isw ? net.mw : net.mb, *v = isw ? net.vw : net.vb;
	const float g = isw ? net.gw[i] : net.gb[i];
	m[i] = b1 * m[i] + (1.f - b1) * g;
	v[i] = b2 * v[i] + (1.f - b2) * g * g;
	const float mh = m[i] / (1.f - powf(b1, (float) t)), vh = v[i] / (1.f - powf(b2, (float) t));
	w[i] -= lr * mh / (sqrtf(vh) + eps);
}

void launch_net_forward(const NetDev &net, const __bf16 *x, int m, const int *m_dev, float *y, __bf16 *h,
                        hipStream_t s) {
	const int blocks = (m + 15) / 16;
	if(blocks > 0) hipLaunchKernelGGL(k_net_fwd, dim3(blocks), dim3(64), 0, s, net, x, m, m_dev, y, h, m);
}

void launch_net_pack(const NetDev &net, hipStream_t s) {
	hipLaunchKernelGGL(k_net_pack, dim3((NN_W + 255) / 256), dim3(256), 0, s, net);
}

// m_pad (a multiple of 32) rows of x (rows >= m zero): forward with stored
// activations, loss, backward, gradients into net.gw / net.gb; loss[m_pad].
void launch_net_loss_grad(const NetDev &net, const __bf16 *x, const float *target, int m, __bf16 *h, float *y,
                          __bf16 *dy, __bf16 *dh, float *loss, hipStream_t s) {
	const int m_pad = (m + 31) & ~31;
	hipLaunchKernelGGL(k_net_fwd, dim3(m_pad / 16), dim3(64), 0, s, net, x, m_pad, (const int *) nullptr, y, h, m_pad);
	hipLaunchKernelGGL(k_net_loss, dim3((m_pad + 255) / 256), dim3(256), 0, s, y, target, m, m_pad, dy, loss);
	hipLaunchKernelGGL(k_net_bwd, dim3(m_pad / 16), dim3(64), 0, s, net, dy, h, m_pad, m_pad, dh);
	int tiles = 0;
	for(int l = 0; l < NN_LAYERS; l++) tiles += (nn_k(l) / 16) * (nn_n(l) / 16);
	hipLaunchKernelGGL(k_net_wgrad, dim3(tiles), dim3(64), 0, s, net, x, h, dy, dh, m_pad, m_pad);
	hipLaunchKernelGGL(k_net_bgrad, dim3((NN_B + 63) / 64), dim3(64), 0, s, net, dy, dh, m_pad, m_pad);
}

void launch_net_adam(const NetDev &net, int step, float lr, hipStream_t s) {
	hipLaunchKernelGGL(k_net_adam, dim3((NN_W + NN_B + 255) / 256), dim3(256), 0, s, net, step, lr);
	launch_net_pack(net, s);
}

// ------------------------------------------------------------------------
// random draws of a learned move

__device__ inline void gauss2(uint32_t u1, uint32_t u2, double &z0, double &z1) {
	const double a = (u1 + 0.5) * 2.3283064365386962890625e-10;  // (0, 1)
	const double b = u01(u2);
	const double rr = sqrt(-2. * log(a));
	double sn, cs;
	sincos(6.283185307179586 * b, &sn, &cs);
	z0 = rr * cs;
	z1 = rr * sn;
}

// vMF sample on S^2 around unit mu (Wood 1994: the cosine by inverse CDF)
__device__ inline void vmf_sample(const float muf[3], double kappa, double u1, double u2, double out[3]) {
	double mu[3] = {muf[0], muf[1], muf[2]};
	const double n = sqrt(mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2]);
	for(int k = 0; k < 3; k++) mu[k] /= n;
	double w;
	if(kappa < 1e-6) w = 2. * u1 - 1.;
	else w = 1. + log(exp(-2. * kappa) + u1 * (-expm1(-2. * kappa))) / kappa;
	w = fmin(1., fmax(-1., w));
	// an orthonormal pair perpendicular to mu
	double e1[3];
	if(fabs(mu[0]) < 0.9) e1[0] = 0., e1[1] = -mu[2], e1[2] = mu[1];
	else e1[0] = mu[2], e1[1] = 0., e1[2] = -mu[0];
	const double n1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
	for(int k = 0; k < 3; k++) e1[k] /= n1;
	const double e2[3] = {mu[1] * e1[2] - mu[2] * e1[1], mu[2] * e1[0] - mu[0] * e1[2], mu[0] * e1[1] - mu[1] * e1[0]};
	double sn, cs;
	sincos(6.283185307179586 * u2, &sn, &cs);
	const double s = sqrt(fmax(0., 1. - w * w));
	for(int k = 0; k < 3; k++) out[k] = w * mu[k] + s * (cs * e1[k] + sn * e2[k]);
	const double no = sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
	for(int k = 0; k < 3; k++) out[k] /= no;
}

// ------------------------------------------------------------------------
// the learned colour pass: one move per non-empty cell of the pass's list

struct CellView {
	const int *rec;
	int M, s0, cnt, cloc, r, sel;
};

__device__ inline CellView cell_view(const DevState &d, int pass, int gi) {
	CellView v;
	v.rec = d.rec + ((size_t) pass * d.total_act + gi) * REC_W;
	v.M = v.rec[58];
	v.s0 = v.rec[59];
	v.cnt = v.rec[60];
	v.cloc = v.rec[61];
	v.r = v.rec[62];
	v.sel = v.rec[63];
	return v;
}

// slot of static neighbour i < M of the cell (the record's 26 ranges)
__device__ inline int rec_slot(const int *rec, int i) {
	int lo = 0, hi = 26;
	while(hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if(rec[32 + mid] <= i) lo = mid;
		else hi = mid;
	}
	return rec[lo] + (i - rec[32 + lo]);
}

// candidate neighbours of the moving slot j: the cell's static neighbours and
// the other members of the cell
struct Cands {
	const double *F;
	int S;
	CellView v;
	int j;
	__device__ int count() const { return v.M + v.cnt - 1; }
	__device__ int slot(int i) const {
		if(i < v.M) return rec_slot(v.rec, i);
		const int o = v.s0 + (i - v.M);
		return o < j ? o : o + 1;
	}
	__device__ void pos(int i, double q[3]) const {
		const int s = slot(i);
		for(int k = 0; k < 3; k++) q[k] = F[(size_t) k * S + s];
	}
	__device__ void ori(int i, double O[9]) const {
		const int s = slot(i);
		for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * S + s];
	}
};

constexpr int PROP_W = NN_PROP_W;

// A: pick the moving member (uniform by rank of particle index, as the
// plain pass) and write the descriptor of the current state.
__global__ __launch_bounds__(64) void k_learn_a(DevState d, uint64_t sweep, int pass) {
	const int gi = blockIdx.x * blockDim.x + threadIdx.x;
	if(gi >= d.pfx[pass * (d.R + 1) + d.R]) return;
	const CellView v = cell_view(d, pass, gi);
	const Rep &rp = d.reps[v.r];
	const double *F = d.f[v.sel];
	const int *PID = d.pid[v.sel];
	const int S = d.stride;
	const u32x4 ra = rng(d.seed, v.r, sweep, (uint32_t) v.cloc, TAG_LEARN_A, (uint32_t) pass, 0);
	const int rank = (int) __umulhi(ra.x, (uint32_t) v.cnt);
	int j = v.s0;
	for(int a = 0; a < v.cnt; a++) {
		int rk = 0;
		for(int b = 0; b < v.cnt; b++) rk += PID[v.s0 + b] < PID[v.s0 + a];
		if(rk == rank) j = v.s0 + a;
	}
	d.learn.slot[gi] = j;
	double p[3], O[9];
	for(int k = 0; k < 3; k++) p[k] = F[(size_t) k * S + j];
	for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * S + j];
	const Cands c{F, S, v, j};
	descriptor(
		p, O, c.count(), [&](int i, double q[3]) { c.pos(i, q); }, [&](int i, double Oq[9]) { c.ori(i, Oq); }, rp.g.box,
		rp.g.inv_box, d.learn.xf + (size_t) gi * NN_IN_PAD);
}

// B: draw the proposal from the forward outputs; reject a move that leaves
// the cell; write the descriptor of the proposed state.
__global__ __launch_bounds__(64) void k_learn_b(DevState d, uint64_t sweep, int pass) {
	const int gi = blockIdx.x * blockDim.x + threadIdx.x;
	if(gi >= d.pfx[pass * (d.R + 1) + d.R]) return;
	const CellView v = cell_view(d, pass, gi);
	const Rep &rp = d.reps[v.r];
	const double *F = d.f[v.sel];
	const int S = d.stride;
	const int j = d.learn.slot[gi];
	Prop pr;
	prop_decode(d.learn.yf + (size_t) gi * NN_OUT_PAD, pr);
	const u32x4 ra = rng(d.seed, v.r, sweep, (uint32_t) v.cloc, TAG_LEARN_A, (uint32_t) pass, 1);
	const u32x4 rb = rng(d.seed, v.r, sweep, (uint32_t) v.cloc, TAG_LEARN_B, (uint32_t) pass, 0);
	double z[4];
	gauss2(ra.x, ra.y, z[0], z[1]);
	gauss2(ra.z, ra.w, z[2], z[3]);
	float dbody[3];
	for(int k = 0; k < 3; k++) dbody[k] = pr.mu[k] + expf(pr.ls[k]) * (float) z[k];
	// mixture component by its weight, then the axis
	const double uc = u01(rb.x);
	int comp = NN_MIX - 1;
	double acc = 0.;
	for(int c = 0; c < NN_MIX; c++) {
		acc += exp((double) pr.lw[c]);
		if(uc < acc) {
			comp = c;
			break;
		}
	}
	double ax[3];
	vmf_sample(pr.dir[comp], exp((double) pr.lk[comp]), u01(rb.y), u01(rb.z), ax);
	const float axf[3] = {(float) ax[0], (float) ax[1], (float) ax[2]};
	const double ang = u01(rb.w) * d.learn.theta_max;
	double p[3], O[9];
	for(int k = 0; k < 3; k++) p[k] = F[(size_t) k * S + j];
	for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * S + j];
	// body -> lab: O^T d; O' = O R(a, angle) (utils.c:186-209)
	double np[3];
	for(int k = 0; k < 3; k++) np[k] = p[k] + (O[k] * dbody[0] + O[3 + k] * dbody[1] + O[6 + k] * dbody[2]);
	const double axd[3] = {axf[0], axf[1], axf[2]};
	double st, ct, R[9], On[9];
	sincos(ang, &st, &ct);
	rotation_matrix(axd, st, ct, R);
	mat_mul3(O, R, On);
	double *out = d.learn.prop + (size_t) gi * PROP_W;
	for(int k = 0; k < 3; k++) out[k] = np[k];
	for(int k = 0; k < 9; k++) out[3 + k] = On[k];
	out[12] = prop_logq(pr, dbody, axf);
	for(int k = 0; k < 3; k++) out[13 + k] = axd[k];
	const int n0 = rp.g.n[0], n1 = rp.g.n[1], n2 = rp.g.n[2];
	const int ix = v.cloc / (n1 * n2), iy = (v.cloc / n2) % n1, iz = v.cloc % n2;
	if(cell_coord(np[0], rp.g.shift[0], rp.g.inv_box[0], n0) != ix ||
	   cell_coord(np[1], rp.g.shift[1], rp.g.inv_box[1], n1) != iy ||
	   cell_coord(np[2], rp.g.shift[2], rp.g.inv_box[2], n2) != iz) {
		d.learn.slot[gi] = -1 - j;  // leaves the cell: rejected, no reverse density needed
		return;
	}
	const Cands c{F, S, v, j};
	descriptor(
		np, On, c.count(), [&](int i, double q[3]) { c.pos(i, q); }, [&](int i, double Oq[9]) { c.ori(i, Oq); }, rp.g.box,
		rp.g.inv_box, d.learn.xr + (size_t) gi * NN_IN_PAD);
}

// C: reverse density, Kern-Frenkel energy change (the reference's pair test),
// Metropolis-Hastings, write the accepted state.
__global__ __launch_bounds__(64) void k_learn_c(DevState d, uint64_t sweep, int pass) {
	const int gi = blockIdx.x * blockDim.x + threadIdx.x;
	if(gi >= d.pfx[pass * (d.R + 1) + d.R]) return;
	const CellView v = cell_view(d, pass, gi);
	const int jj = d.learn.slot[gi];
	const int cslot = gi & (NSLOT - 1);
	atomicAdd(ctr_at(d.ctr, v.r, C_TRY_ROTO, cslot), 1ull);
	atomicAdd(ctr_at(d.ctr, v.r, C_TRY_LEARN, cslot), 1ull);
	if(jj < 0) return;
	const int j = jj;
	const Rep &rp = d.reps[v.r];
	double *F = d.f[v.sel];
	const int S = d.stride;
	const double *pp = d.learn.prop + (size_t) gi * PROP_W;
	double p[3], O[9], np[3], On[9];
	for(int k = 0; k < 3; k++) p[k] = F[(size_t) k * S + j];
	for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * S + j];
	for(int k = 0; k < 3; k++) np[k] = pp[k];
	for(int k = 0; k < 9; k++) On[k] = pp[3 + k];
	// reverse: displacement p - p' in the body frame of O', axis -a
	float rbody[3];
	for(int k = 0; k < 3; k++)
		rbody[k] = (float) (On[3 * k] * (p[0] - np[0]) + On[3 * k + 1] * (p[1] - np[1]) + On[3 * k + 2] * (p[2] - np[2]));
	const float rax[3] = {(float) -pp[13], (float) -pp[14], (float) -pp[15]};
	Prop pr;
	prop_decode(d.learn.yr + (size_t) gi * NN_OUT_PAD, pr);
	const double log_ratio = (double) prop_logq(pr, rbody, rax) - pp[12];
	// energy change over the static neighbours and the other members
	Trial t;
	t.px = p[0];
	t.py = p[1];
	t.pz = p[2];
	t.nx = np[0];
	t.ny = np[1];
	t.nz = np[2];
	for(int k = 0; k < 3; k++) {
		(&t.Lx)[k] = rp.g.box[k];
		(&t.iLx)[k] = rp.g.inv_box[k];
	}
	t.rc2 = rp.rc2;
	t.cosmax = rp.cosmax;
	t.c2 = -1.;
	double Po[4][3], Pn[4][3];
	kf_patches(O, Po);
	kf_patches(On, Pn);
	const Cands c{F, S, v, j};
	int bo = 0, bn = 0;
	bool ov = false;
	for(int i = 0; i < c.count() && !ov; i++) {
		double q[3];
		c.pos(i, q);
		if(overlaps_new(t, q[0], q[1], q[2])) {
			ov = true;
			break;
		}
		bonds_old_new_exact(t, &Po[0][0], &Pn[0][0], q[0], q[1], q[2], [&](double *Oq) { c.ori(i, Oq); }, bo, bn);
	}
	if(ov) return;
	const int db = bn - bo;
	const double dE = (double) (-db);
	const u32x4 rc = rng(d.seed, v.r, sweep, (uint32_t) v.cloc, TAG_LEARN_B, (uint32_t) pass, 1);
	const double la = -dE / rp.T + log_ratio;
	if(la >= 0. || u01(rc.x) < exp(la)) {
		for(int k = 0; k < 3; k++) F[(size_t) k * S + j] = np[k];
		for(int k = 0; k < 9; k++) F[(size_t) (3 + k) * S + j] = On[k];
		atomicAdd(ctr_at(d.ctr, v.r, C_ACC_ROTO, cslot), 1ull);
		atomicAdd(ctr_at(d.ctr, v.r, C_ACC_LEARN, cslot), 1ull);
		if(db) atomicAdd(ctr_at(d.ctr, v.r, C_DBONDS, cslot), (unsigned long long) (long long) db);
	}
}

void launch_learned_pass(const DevState &d, const NetDev &net, uint64_t sweep, int pass, hipStream_t s) {
	const int B = 64;
	const unsigned g = (unsigned) ((d.total_act + B - 1) / B);
	const int *count = d.pfx + pass * (d.R + 1) + d.R;
	hipLaunchKernelGGL(k_learn_a, dim3(g), dim3(B), 0, s, d, sweep, pass);
	launch_net_forward(net, d.learn.xf, d.total_act, count, d.learn.yf, nullptr, s);
	hipLaunchKernelGGL(k_learn_b, dim3(g), dim3(B), 0, s, d, sweep, pass);
	launch_net_forward(net, d.learn.xr, d.total_act, count, d.learn.yr, nullptr, s);
	hipLaunchKernelGGL(k_learn_c, dim3(g), dim3(B), 0, s, d, sweep, pass);
}

// ------------------------------------------------------------------------
// training samples: a random particle of the current state, displaced and
// rotated at random (the "current" state of a move), its descriptor there,
// and as target the move back: displacement in the perturbed body frame and
// the rotation axis (-a; the angle is uniform in both directions).

__global__ __launch_bounds__(128) void k_train_samples(DevState d, uint64_t seed, uint64_t step, int m, int m_pad, double pos_noise,
                                double theta_max, __bf16 *x, float *target) {
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if(b >= m_pad) return;
	if(b >= m) {
		for(int e = 0; e < NN_IN_PAD; e++) x[(size_t) b * NN_IN_PAD + e] = (__bf16) 0.f;
		for(int e = 0; e < 6; e++) target[(size_t) b * 6 + e] = 0.f;
		return;
	}
	const uint32_t k0 = (uint32_t) seed, k1 = (uint32_t) (seed >> 32);
	const u32x4 r0 = philox(u32x4{(uint32_t) b, 0xA5A50000u, (uint32_t) step, (uint32_t) (step >> 32)}, k0, k1);
	const u32x4 r1 = philox(u32x4{(uint32_t) b, 0xA5A50001u, (uint32_t) step, (uint32_t) (step >> 32)}, k0, k1);
	const u32x4 r2 = philox(u32x4{(uint32_t) b, 0xA5A50002u, (uint32_t) step, (uint32_t) (step >> 32)}, k0, k1);
	// a uniform slot of the whole context: a uniform particle of a replica
	// chosen with probability proportional to its size
	const int j = (int) __umulhi(r0.x, (uint32_t) d.Ntot);
	const int rr = d.prep[j];
	const Rep &rp = d.reps[rr];
	const double *F = d.f[rp.sel];
	const int S = d.stride;
	double p[3], O[9];
	for(int k = 0; k < 3; k++) p[k] = F[(size_t) k * S + j];
	for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * S + j];
	// its cell on the current grid (particles never leave their cells between
	// rebuilds) and the 27 cells around it; j itself is left out
	const int n0 = rp.g.n[0], n1 = rp.g.n[1], n2 = rp.g.n[2];
	const int ix = cell_coord(p[0], rp.g.shift[0], rp.g.inv_box[0], n0);
	const int iy = cell_coord(p[1], rp.g.shift[1], rp.g.inv_box[1], n1);
	const int iz = cell_coord(p[2], rp.g.shift[2], rp.g.inv_box[2], n2);
	int st[27], cn[27], tot = 0, self = -1;
	for(int nb = 0; nb < 27; nb++) {
		const int cx = wrapi(ix + nb / 9 - 1, n0), cy = wrapi(iy + (nb / 3) % 3 - 1, n1), cz = wrapi(iz + nb % 3 - 1, n2);
		const int cc = rp.cell_off + (cx * n1 + cy) * n2 + cz;
		st[nb] = d.start[cc];
		cn[nb] = d.start[cc + 1] - st[nb];
		if(j >= st[nb] && j < st[nb] + cn[nb]) self = tot + (j - st[nb]);
		tot += cn[nb];
	}
	const int n = self >= 0 ? tot - 1 : tot;
	auto slot_of = [&](int i) {
		if(self >= 0 && i >= self) i++;
		int nb = 0;
		while(nb < 26 && i >= cn[nb]) {
			i -= cn[nb];
			nb++;
		}
		return st[nb] + i;
	};
	// the perturbed state: Gaussian displacement, rotation by an angle uniform
	// in [0, theta_max) about a uniform axis
	double z[6];
	gauss2(r0.y, r0.z, z[0], z[1]);
	gauss2(r0.w, r1.x, z[2], z[3]);
	gauss2(r1.y, r1.z, z[4], z[5]);
	double pt[3];
	for(int k = 0; k < 3; k++) pt[k] = p[k] + pos_noise * z[k];
	double ax[3] = {z[3], z[4], z[5]};
	const double an = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
	for(int k = 0; k < 3; k++) ax[k] /= an;
	const double ang = u01(r1.w) * theta_max;
	(void) r2;
	double sn, cs, R[9], Ot[9];
	sincos(ang, &sn, &cs);
	rotation_matrix(ax, sn, cs, R);
	mat_mul3(O, R, Ot);
	auto pos = [&](int i, double q[3]) {
		const int s2 = slot_of(i);
		for(int k = 0; k < 3; k++) q[k] = F[(size_t) k * S + s2];
	};
	auto ori = [&](int i, double Oq[9]) {
		const int s2 = slot_of(i);
		for(int k = 0; k < 9; k++) Oq[k] = F[(size_t) (3 + k) * S + s2];
	};
	descriptor(pt, Ot, n, pos, ori, rp.g.box, rp.g.inv_box, x + (size_t) b * NN_IN_PAD);
	float *t = target + (size_t) b * 6;
	for(int k = 0; k < 3; k++)
		t[k] = (float) (Ot[3 * k] * (p[0] - pt[0]) + Ot[3 * k + 1] * (p[1] - pt[1]) + Ot[3 * k + 2] * (p[2] - pt[2]));
	for(int k = 0; k < 3; k++) t[3 + k] = (float) -ax[k];
}

void launch_train_samples(const DevState &d, uint64_t seed, uint64_t step, int m, double pos_noise, double theta_max,
                          __bf16 *x, float *target, hipStream_t s) {
	const int m_pad = (m + 31) & ~31;
	hipLaunchKernelGGL(k_train_samples, dim3((m_pad + 127) / 128), dim3(128), 0, s, d, seed, step, m, m_pad, pos_noise,
	                   theta_max, x, target);
}

// Descriptors of particles idx[t] of a plain configuration, every other
// particle a candidate (the C ABI's pp_net_descriptor).
__global__ void k_descriptor(int n, const double *r, const double *O, Grid g, int m, const int *idx, __bf16 *x) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if(t >= m) return;
	const int i = idx[t];
	double p[3], Oi[9];
	for(int k = 0; k < 3; k++) p[k] = r[3 * (size_t) i + k];
	for(int k = 0; k < 9; k++) Oi[k] = O[9 * (size_t) i + k];
	auto other = [&](int c) { return c < i ? c : c + 1; };
	descriptor(
		p, Oi, n - 1,
		[&](int c, double q[3]) {
			for(int k = 0; k < 3; k++) q[k] = r[3 * (size_t) other(c) + k];
		},
		[&](int c, double Oq[9]) {
			for(int k = 0; k < 9; k++) Oq[k] = O[9 * (size_t) other(c) + k];
		},
		g.box, g.inv_box, x + (size_t) t * NN_IN_PAD);
}

void launch_descriptor(int n, const double *r, const double *O, const Grid &g, int m, const int *idx, __bf16 *x,
                       hipStream_t s) {
	if(m > 0) hipLaunchKernelGGL(k_descriptor, dim3((m + 63) / 64), dim3(64), 0, s, n, r, O, g, m, idx, x);
}

}  // namespace ppd
