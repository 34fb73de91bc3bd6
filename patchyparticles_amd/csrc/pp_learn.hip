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

// Descriptor of a particle at (p, O) among n candidate neighbours (pos(i, q)
// and ori(i, Oq) give candidate i): the NN_NB nearest within NN_RCUT (minimum
// image; squared distances compared in single precision, ties by candidate
// order), nearest first, each as its separation in the body frame of O and
// the quaternion of Oq O^T; zero-padded to NN_IN_PAD.  One wavefront per
// particle, in two phases with a workgroup barrier between them:
//   desc_collect: zero the row; the in-range candidates, in candidate order,
//                 into the wave's LDS list (at most DESC_CAND are kept);
//   desc_fill:    each listed candidate's rank by counting; the NN_NB first
//                 write their features.
constexpr int DESC_CAND = 128;

struct DescLds {
	float f[DESC_CAND];  // squared distance of an in-range candidate
	int i[DESC_CAND];    // its candidate index
};

// T: lanes per particle (64: a wavefront, 32: half of one); tl = lane % T.
template <int T, class Pos>
__device__ int desc_collect(const double p[3], int n, Pos pos, const double L[3], const double iL[3], DescLds &S,
                            __bf16 *row, int tl) {
	for(int e = tl; e < NN_IN_PAD / 2; e += T) {
		row[2 * e] = (__bf16) 0.f;
		row[2 * e + 1] = (__bf16) 0.f;
	}
	const double rc2 = NN_RCUT * NN_RCUT;
	const int off = (threadIdx.x & 63) & ~(T - 1);
	const unsigned long long tm = T == 64 ? ~0ull : ((1ull << T) - 1ull);
	int nv = 0;
	for(int base = 0; base < n; base += T) {
		const int i = base + tl;
		bool ok = false;
		float f = 0.f;
		if(i < n) {
			double q[3];
			pos(i, q);
			const double dx = kf_min_image(q[0] - p[0], L[0], iL[0]);
			const double dy = kf_min_image(q[1] - p[1], L[1], iL[1]);
			const double dz = kf_min_image(q[2] - p[2], L[2], iL[2]);
			const double r2 = dx * dx + dy * dy + dz * dz;
			ok = r2 < rc2;
			f = (float) r2;
		}
		const unsigned long long m = (__ballot(ok) >> off) & tm;
		const int at = nv + __popcll(m & ((1ull << tl) - 1ull));
		if(ok && at < DESC_CAND) {
			S.f[at] = f;
			S.i[at] = i;
		}
		nv += __popcll(m);
	}
	return nv < DESC_CAND ? nv : DESC_CAND;
}

template <int T, class Pos, class Ori>
__device__ void desc_fill(const double p[3], const double O[9], int nv, Pos pos, Ori ori, const double L[3],
                          const double iL[3], const DescLds &S, __bf16 *row, int tl) {
	for(int e = tl; e < nv; e += T) {
		const float f = S.f[e];
		const int i = S.i[e];
		int rank = 0;
		for(int k = 0; k < nv; k++) {
			const float g = S.f[k];
			rank += (g < f) || (g == f && S.i[k] < i);
		}
		if(rank >= NN_NB) continue;
		double q[3], Oq[9];
		pos(i, q);
		ori(i, Oq);
		double d[3];
		for(int k = 0; k < 3; k++) d[k] = kf_min_image(q[k] - p[k], L[k], iL[k]);
		float v[NN_FEAT];
		for(int k = 0; k < 3; k++) v[k] = (float) (O[3 * k] * d[0] + O[3 * k + 1] * d[1] + O[3 * k + 2] * d[2]);
		double M[9];
		for(int a = 0; a < 3; a++)
			for(int b = 0; b < 3; b++) M[3 * a + b] = Oq[3 * a] * O[3 * b] + Oq[3 * a + 1] * O[3 * b + 1] + Oq[3 * a + 2] * O[3 * b + 2];
		quat_of(M, v + 3);
		for(int k = 0; k < NN_FEAT; k++) row[rank * NN_FEAT + k] = (__bf16) v[k];
	}
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

// Bias gradients: one workgroup per bias, a tree sum over the rows.
__global__ __launch_bounds__(256) void k_net_bgrad(NetDev net, const __bf16 *dy, const __bf16 *dh, int m_pad, int hs) {
	__shared__ float part[256];
	const int i = blockIdx.x;
	int l = 0;
	while(l + 1 < NN_LAYERS && i >= nn_boff(l + 1)) l++;
	const int n = i - nn_boff(l);
	const __bf16 *Z = l == NN_LAYERS - 1 ? dy : dh + (size_t) l * hs * NN_H;
	const int ldz = l == NN_LAYERS - 1 ? NN_OUT_PAD : NN_H;
	float s = 0.f;
	for(int row = threadIdx.x; row < m_pad; row += 256) s += (float) Z[(size_t) row * ldz + n];
	part[threadIdx.x] = s;
	__syncthreads();
	for(int o = 128; o > 0; o >>= 1) {
		if((int) threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
		__syncthreads();
	}
	if(threadIdx.x == 0) net.gb[i] = part[0];
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
	hipLaunchKernelGGL(k_net_bgrad, dim3(NN_B), dim3(256), 0, s, net, dy, dh, m_pad, m_pad);
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

// Single-precision draws of a learned move.  The move applied is the one the
// density is evaluated at (body-frame displacement and axis as floats), so
// the precision of the transforms only shapes the proposal, not the
// acceptance rule.
__device__ inline void gauss2f(uint32_t u1, uint32_t u2, float &z0, float &z1) {
	const float a = ((float) (u1 >> 8) + 0.5f) * 5.9604644775390625e-8f;  // (0, 1)
	const float b = (float) (u2 >> 8) * 5.9604644775390625e-8f;
	const float rr = sqrtf(-2.f * logf(a));
	float sn, cs;
	sincosf(6.2831853f * b, &sn, &cs);
	z0 = rr * cs;
	z1 = rr * sn;
}

// vMF sample on S^2 around unit mu (Wood 1994: the cosine by inverse CDF)
__device__ inline void vmf_sample(const float mu[3], float kappa, float u1, float u2, float out[3]) {
	float w;
	if(kappa < 1e-4f) w = 2.f * u1 - 1.f;
	else w = 1.f + logf(expf(-2.f * kappa) + u1 * (-expm1f(-2.f * kappa))) / kappa;
	w = fminf(1.f, fmaxf(-1.f, w));
	// an orthonormal pair perpendicular to mu
	float e1[3];
	if(fabsf(mu[0]) < 0.9f) e1[0] = 0.f, e1[1] = -mu[2], e1[2] = mu[1];
	else e1[0] = mu[2], e1[1] = 0.f, e1[2] = -mu[0];
	const float n1 = sqrtf(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
	for(int k = 0; k < 3; k++) e1[k] /= n1;
	const float e2[3] = {mu[1] * e1[2] - mu[2] * e1[1], mu[2] * e1[0] - mu[0] * e1[2], mu[0] * e1[1] - mu[1] * e1[0]};
	float sn, cs;
	sincosf(6.2831853f * u2, &sn, &cs);
	const float sw = sqrtf(fmaxf(0.f, 1.f - w * w));
	for(int k = 0; k < 3; k++) out[k] = w * mu[k] + sw * (cs * e1[k] + sn * e2[k]);
	const float no = sqrtf(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
	for(int k = 0; k < 3; k++) out[k] /= no;
}

__device__ inline float u01f(uint32_t u) { return (float) (u >> 8) * 5.9604644775390625e-8f; }  // [0, 1)

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

// slot of static neighbour i < M of the cell (the record's 26 ranges; rec may
// be the global record or a copy of it in LDS)
__device__ inline int rec_slot(const int *rec, int i) {
	int lo = 0, hi = 26;
	while(hi - lo > 1) {
		const int mid = (lo + hi) >> 1;
		if(rec[32 + mid] <= i) lo = mid;
		else hi = mid;
	}
	return rec[lo] + (i - rec[32 + lo]);
}

// Candidates of one wave's cell, staged in LDS: their slots and positions
// (the first CAND_MAX; the rest are read from the record and global memory).
constexpr int CAND_MAX = 128;
struct CandLds {
	double x[CAND_MAX], y[CAND_MAX], z[CAND_MAX];
	int slot[CAND_MAX];
	int rec[REC_W];  // the cell's record
};

// candidate neighbours of the moving slot j: the cell's static neighbours and
// the other members of the cell
struct Cands {
	const double *F;
	int S;
	CellView v;
	int j;
	const CandLds *L = nullptr;  // staged slots / positions (filled by stage())
	__device__ int count() const { return v.M + v.cnt - 1; }
	__device__ int slot_of(int i) const {
		if(i < v.M) return rec_slot(L ? L->rec : v.rec, i);
		const int o = v.s0 + (i - v.M);
		return o < j ? o : o + 1;
	}
	__device__ int slot(int i) const { return (L && i < CAND_MAX) ? L->slot[i] : slot_of(i); }
	__device__ void pos(int i, double q[3]) const {
		if(L && i < CAND_MAX) {
			q[0] = L->x[i];
			q[1] = L->y[i];
			q[2] = L->z[i];
			return;
		}
		const int s = slot_of(i);
		for(int k = 0; k < 3; k++) q[k] = F[(size_t) k * S + s];
	}
	// lane-strided: candidate i is staged by lane i % T of the team, the
	// lane that reads it first (desc_collect, the energy loop)
	template <int T>
	__device__ void stage(CandLds &W, int tl) const {
		for(int i = tl; i < count() && i < CAND_MAX; i += T) {
			const int s = slot_of(i);
			W.slot[i] = s;
			W.x[i] = F[s];
			W.y[i] = F[(size_t) S + s];
			W.z[i] = F[(size_t) 2 * S + s];
		}
	}
	__device__ void ori(int i, double O[9]) const {
		const int s = slot(i);
		for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * S + s];
	}
};

constexpr int PROP_W = NN_PROP_W;  // p'[3], O'[9], log q forward, axis[3]
constexpr int LP_ROWS = 16;        // rows of the MFMA tile = cells of one workgroup
constexpr int LP_TEAM = 32;        // lanes per cell: two cells per wavefront
constexpr int LP_WAVES = LP_ROWS * LP_TEAM / 64;

// LDS of one learned-pass workgroup
struct LpLds {
	__bf16 x[LP_ROWS * ACT_LD];       // descriptors
	__bf16 act[2][LP_ROWS * ACT_LD];  // hidden activations
	float y[LP_ROWS * NN_OUT_PAD];    // network outputs
	DescLds desc[LP_ROWS];
	double prop[LP_ROWS][PROP_W];     // the drawn proposal of each cell
	CandLds cand[LP_ROWS];
};

// The network on the workgroup's 16 rows, S.x -> S.y; wave w computes the
// 16-column tiles w, w + LP_WAVES, ... of each layer; a barrier after each.
__device__ void lp_forward(const NetDev &net, LpLds &S, int wave, int lane) {
	const int r = lane & 15, g = lane >> 4;
	int cur = 0;
	for(int l = 0; l < NN_LAYERS; l++) {
		const int nkb = nn_k(l) / 32, nnt = nn_n(l) / 16;
		const __bf16 *A = l == 0 ? S.x : S.act[cur];
		for(int nt = wave; nt < nnt; nt += LP_WAVES) {
			f32x4 acc = {0.f, 0.f, 0.f, 0.f};
			for(int kb = 0; kb < nkb; kb++) {
				const bf16x8 a = *(const bf16x8 *) (A + r * ACT_LD + kb * 32 + g * 8);
				const bf16x8 b = *(const bf16x8 *) (net.frag + frag_at(l, kb, nt, lane));
				acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
			}
			const int col = nt * 16 + r;
			const float bias = net.b[nn_boff(l) + col];
			for(int j = 0; j < 4; j++) {
				const int row = g * 4 + j;
				const float v = acc[j] + bias;
				if(l < NN_LAYERS - 1) S.act[cur ^ 1][row * ACT_LD + col] = (__bf16) fmaxf(v, 0.f);
				else S.y[row * NN_OUT_PAD + col] = v;
			}
		}
		__syncthreads();
		cur ^= 1;
	}
}

// The whole learned pass in one kernel: a workgroup takes 16 cells of the
// pass's list (the rows of one MFMA tile), 32 lanes per cell:
//   A  pick the moving member (uniform by rank of particle index, as the plain
//      pass), descriptor of the current state;      network forward (MFMA)
//   B  draw the proposal from the outputs; a move that leaves the cell is
//      rejected; descriptor of the proposed state;  network forward (MFMA)
//   C  reverse density, Kern-Frenkel energy change over the cell's
//      candidates (the reference's pair test), Metropolis-Hastings.
__global__ __launch_bounds__(64 * LP_WAVES) void k_learn_pass(DevState d, NetDev net, uint64_t sweep, int pass) {
	__shared__ __attribute__((aligned(16))) LpLds S;
	constexpr int T = LP_TEAM;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int tl = lane & (T - 1), cell = threadIdx.x / T;
	const int count = d.pfx[pass * (d.R + 1) + d.R];
	const int g0 = blockIdx.x * LP_ROWS;
	if(g0 >= count) return;  // the whole workgroup: before any barrier
	const int gi = g0 + cell;
	bool live = gi < count;
	CellView v{};
	const double *F = nullptr;
	const int SS = d.stride;
	int j = 0, nv = 0;
	__bf16 *row = S.x + cell * ACT_LD;
	DescLds &D = S.desc[cell];
	CandLds &CL = S.cand[cell];
	double L[3], iL[3];
	if(live) {
		v = cell_view(d, pass, gi);
		// a cell of more members than a team holds (one per lane) is not
		// moved and is reported; its threads stay for the barriers below
		if(v.cnt > CELL_CAP || v.cnt > T) {
			if(tl == 0) atomicOr(d.err, (int) ERR_CELL_CAP);
			live = false;
		}
		else
			for(int e = tl; e < REC_W; e += T) CL.rec[e] = v.rec[e];
	}
	__syncthreads();

	// ---- A
	if(live) {
		const Rep &rp = d.reps[v.r];
		for(int k = 0; k < 3; k++) {
			L[k] = rp.g.box[k];
			iL[k] = rp.g.inv_box[k];
		}
		F = pick_buf(d.f, v.sel);
		const int *PID = pick_buf(d.pid, v.sel);
		const u32x4 ra = rng(d.seed, v.r, sweep, (uint32_t) v.cloc, TAG_LEARN_A, (uint32_t) pass, 0);
		const int target = (int) __umulhi(ra.x, (uint32_t) v.cnt);
		bool hit = false;
		if(tl < v.cnt) {
			const int me = PID[v.s0 + tl];
			int rk = 0;
			for(int b = 0; b < v.cnt; b++) rk += PID[v.s0 + b] < me;
			hit = rk == target;
		}
		const unsigned long long hm = (__ballot(hit) >> (lane & ~(T - 1))) & ((1ull << T) - 1ull);
		j = v.s0 + (int) __builtin_ctzll(hm);
		const Cands c{F, SS, v, j, &CL};
		c.stage<T>(CL, tl);
		double p[3];
		for(int k = 0; k < 3; k++) p[k] = F[(size_t) k * SS + j];
		nv = desc_collect<T>(p, c.count(), [&](int i, double q[3]) { c.pos(i, q); }, L, iL, D, row, tl);
	}
	else
		for(int e = tl; e < NN_IN_PAD / 2; e += T) {
			row[2 * e] = (__bf16) 0.f;
			row[2 * e + 1] = (__bf16) 0.f;
		}
	__syncthreads();
	if(live) {
		const Cands c{F, SS, v, j, &CL};
		double p[3], O[9];
		for(int k = 0; k < 3; k++) p[k] = F[(size_t) k * SS + j];
		for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * SS + j];
		desc_fill<T>(
			p, O, nv, [&](int i, double q[3]) { c.pos(i, q); }, [&](int i, double Oq[9]) { c.ori(i, Oq); }, L, iL, D, row,
			tl);
	}
	__syncthreads();
	lp_forward(net, S, wave, lane);

	// ---- B
	bool inside = false;
	if(live) {
		const Rep &rp = d.reps[v.r];
		Prop pr;
		prop_decode(S.y + cell * NN_OUT_PAD, pr);
		const u32x4 ra = rng(d.seed, v.r, sweep, (uint32_t) v.cloc, TAG_LEARN_A, (uint32_t) pass, 1);
		const u32x4 rb = rng(d.seed, v.r, sweep, (uint32_t) v.cloc, TAG_LEARN_B, (uint32_t) pass, 0);
		float z[4];
		gauss2f(ra.x, ra.y, z[0], z[1]);
		gauss2f(ra.z, ra.w, z[2], z[3]);
		float dbody[3];
		for(int k = 0; k < 3; k++) dbody[k] = pr.mu[k] + expf(pr.ls[k]) * z[k];
		// mixture component by its weight, then the axis
		const float uc = u01f(rb.x);
		int comp = NN_MIX - 1;
		float acc = 0.f;
		for(int c = 0; c < NN_MIX; c++) {
			acc += expf(pr.lw[c]);
			if(uc < acc) {
				comp = c;
				break;
			}
		}
		float dir[3], kap = expf(pr.lk[0]);
		for(int k = 0; k < 3; k++) dir[k] = pr.dir[0][k];
		for(int c = 1; c < NN_MIX; c++)
			if(c == comp) {
				for(int k = 0; k < 3; k++) dir[k] = pr.dir[c][k];
				kap = expf(pr.lk[c]);
			}
		float axf[3];
		vmf_sample(dir, kap, u01f(rb.y), u01f(rb.z), axf);
		const double ang = u01(rb.w) * d.learn.theta_max;
		double p[3], O[9];
		for(int k = 0; k < 3; k++) p[k] = F[(size_t) k * SS + j];
		for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * SS + j];
		// body -> lab: O^T d; O' = O R(a, angle) (utils.c:186-209)
		double np[3];
		for(int k = 0; k < 3; k++) np[k] = p[k] + (O[k] * dbody[0] + O[3 + k] * dbody[1] + O[6 + k] * dbody[2]);
		const double axd[3] = {axf[0], axf[1], axf[2]};
		double st, ct, R[9], On[9];
		sincos(ang, &st, &ct);
		rotation_matrix(axd, st, ct, R);
		mat_mul3(O, R, On);
		double *out = S.prop[cell];
		if(tl == 0) {
			for(int k = 0; k < 3; k++) out[k] = np[k];
			for(int k = 0; k < 9; k++) out[3 + k] = On[k];
			out[12] = prop_logq(pr, dbody, axf);
			for(int k = 0; k < 3; k++) out[13 + k] = axd[k];
		}
		const int n0 = rp.g.n[0], n1 = rp.g.n[1], n2 = rp.g.n[2];
		const int ix = v.cloc / (n1 * n2), iy = (v.cloc / n2) % n1, iz = v.cloc % n2;
		inside = cell_coord(np[0], rp.g.shift[0], rp.g.inv_box[0], n0) == ix &&
		         cell_coord(np[1], rp.g.shift[1], rp.g.inv_box[1], n1) == iy &&
		         cell_coord(np[2], rp.g.shift[2], rp.g.inv_box[2], n2) == iz;
		if(inside) {
			const Cands c{F, SS, v, j, &CL};
			nv = desc_collect<T>(np, c.count(), [&](int i, double q[3]) { c.pos(i, q); }, L, iL, D, row, tl);
		}
	}
	__syncthreads();
	if(live && inside) {
		const Cands c{F, SS, v, j, &CL};
		const double *pp = S.prop[cell];
		double np[3], On[9];
		for(int k = 0; k < 3; k++) np[k] = pp[k];
		for(int k = 0; k < 9; k++) On[k] = pp[3 + k];
		desc_fill<T>(
			np, On, nv, [&](int i, double q[3]) { c.pos(i, q); }, [&](int i, double Oq[9]) { c.ori(i, Oq); }, L, iL, D, row,
			tl);
	}
	__syncthreads();
	lp_forward(net, S, wave, lane);

	// ---- C (no barrier below: teams leave independently)
	if(!live) return;
	const int cslot = gi & (NSLOT - 1);
	if(tl == 0) {
		atomicAdd(ctr_at(d.ctr, v.r, C_TRY_ROTO, cslot), 1ull);
		atomicAdd(ctr_at(d.ctr, v.r, C_TRY_LEARN, cslot), 1ull);
	}
	if(!inside) return;
	const Rep &rp = d.reps[v.r];
	double *Fw = pick_buf(d.f, v.sel);
	const double *pp = S.prop[cell];
	double p[3], O[9], np[3], On[9];
	for(int k = 0; k < 3; k++) p[k] = Fw[(size_t) k * SS + j];
	for(int k = 0; k < 9; k++) O[k] = Fw[(size_t) (3 + k) * SS + j];
	for(int k = 0; k < 3; k++) np[k] = pp[k];
	for(int k = 0; k < 9; k++) On[k] = pp[3 + k];
	// reverse: displacement p - p' in the body frame of O', axis -a
	float rbody[3];
	for(int k = 0; k < 3; k++)
		rbody[k] = (float) (On[3 * k] * (p[0] - np[0]) + On[3 * k + 1] * (p[1] - np[1]) + On[3 * k + 2] * (p[2] - np[2]));
	const float rax[3] = {(float) -pp[13], (float) -pp[14], (float) -pp[15]};
	Prop pr;
	prop_decode(S.y + cell * NN_OUT_PAD, pr);
	const double log_ratio = (double) prop_logq(pr, rbody, rax) - pp[12];
	// energy change over the static neighbours and the other members, one
	// candidate per lane of the team
	Trial t;
	t.px = p[0];
	t.py = p[1];
	t.pz = p[2];
	t.nx = np[0];
	t.ny = np[1];
	t.nz = np[2];
	t.Lx = L[0];
	t.Ly = L[1];
	t.Lz = L[2];
	t.iLx = iL[0];
	t.iLy = iL[1];
	t.iLz = iL[2];
	t.rc2 = rp.rc2;
	t.cosmax = rp.cosmax;
	t.c2 = -1.;
	double Po[4][3], Pn[4][3];
	kf_patches(O, Po);
	kf_patches(On, Pn);
	const Cands c{Fw, SS, v, j, &CL};
	const int n = c.count();
	int bo = 0, bn = 0;
	bool ov = false;
	for(int base = 0; base < n; base += T) {
		const int i = base + tl;
		if(i < n && !ov) {
			double q[3];
			c.pos(i, q);
			if(overlaps_new(t, q[0], q[1], q[2])) ov = true;
			else bonds_old_new_exact(t, &Po[0][0], &Pn[0][0], q[0], q[1], q[2], [&](double *Oq) { c.ori(i, Oq); }, bo, bn);
		}
		if(team_any<T>(ov)) break;
	}
	if(team_any<T>(ov)) return;
	const int db = team_sum<T>(bn) - team_sum<T>(bo);
	if(tl != 0) return;
	const double dE = (double) (-db);
	const u32x4 rc = rng(d.seed, v.r, sweep, (uint32_t) v.cloc, TAG_LEARN_B, (uint32_t) pass, 1);
	const double la = -dE / rp.T + log_ratio;
	if(la >= 0. || u01(rc.x) < exp(la)) {
		for(int k = 0; k < 3; k++) Fw[(size_t) k * SS + j] = np[k];
		for(int k = 0; k < 9; k++) Fw[(size_t) (3 + k) * SS + j] = On[k];
		atomicAdd(ctr_at(d.ctr, v.r, C_ACC_ROTO, cslot), 1ull);
		atomicAdd(ctr_at(d.ctr, v.r, C_ACC_LEARN, cslot), 1ull);
		if(db) atomicAdd(ctr_at(d.ctr, v.r, C_DBONDS, cslot), (unsigned long long) (long long) db);
	}
}

void launch_learned_pass(const DevState &d, const NetDev &net, uint64_t sweep, int pass, hipStream_t s) {
	const unsigned g = (unsigned) ((d.total_act + LP_ROWS - 1) / LP_ROWS);
	if(g > 0) hipLaunchKernelGGL(k_learn_pass, dim3(g), dim3(64 * LP_WAVES), 0, s, d, net, sweep, pass);
}

// Descriptors of particles idx[t] of a plain configuration, every other
// particle a candidate (the C ABI's pp_net_descriptor); one wave per particle.
__global__ __launch_bounds__(64) void k_descriptor(int n, const double *r, const double *O, Grid g, int m,
                                                   const int *idx, __bf16 *x) {
	__shared__ DescLds D;
	const int t = blockIdx.x, lane = threadIdx.x;
	if(t >= m) return;
	const int i = idx[t];
	double p[3], Oi[9];
	for(int k = 0; k < 3; k++) p[k] = r[3 * (size_t) i + k];
	for(int k = 0; k < 9; k++) Oi[k] = O[9 * (size_t) i + k];
	auto other = [&](int c) { return c < i ? c : c + 1; };
	auto pos = [&](int c, double q[3]) {
		for(int k = 0; k < 3; k++) q[k] = r[3 * (size_t) other(c) + k];
	};
	auto ori = [&](int c, double Oq[9]) {
		for(int k = 0; k < 9; k++) Oq[k] = O[9 * (size_t) other(c) + k];
	};
	__bf16 *row = x + (size_t) t * NN_IN_PAD;
	const int nv = desc_collect<64>(p, n - 1, pos, g.box, g.inv_box, D, row, lane);
	__syncthreads();
	desc_fill<64>(p, Oi, nv, pos, ori, g.box, g.inv_box, D, row, lane);
}

void launch_descriptor(int n, const double *r, const double *O, const Grid &g, int m, const int *idx, __bf16 *x,
                       hipStream_t s) {
	if(m > 0) hipLaunchKernelGGL(k_descriptor, dim3(m), dim3(64), 0, s, n, r, O, g, m, idx, x);
}

// ------------------------------------------------------------------------
// training samples: a random particle of the current state, displaced and
// rotated at random (the "current" state of a move), its descriptor there,
// and as target the move back: displacement in the perturbed body frame and
// the rotation axis (-a; the angle is uniform in both directions).  One wave
// per sample.

constexpr int TS_WAVES = 4;

__global__ __launch_bounds__(64 * TS_WAVES) void k_train_samples(DevState d, uint64_t seed, uint64_t step, int m,
                                                                 int m_pad, double pos_noise, double theta_max,
                                                                 __bf16 *x, float *target) {
	__shared__ DescLds Ds[TS_WAVES];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int b = blockIdx.x * TS_WAVES + wave;
	__bf16 *row = x + (size_t) b * NN_IN_PAD;
	const bool live = b < m;
	int nv = 0;
	double pt[3], Ot[9];
	// the slot's cell neighbourhood (27 cells, the particle itself left out)
	int st[27], cn[27], self = -1, j = 0, tot = 0;
	const double *F = nullptr;
	const int S = d.stride;
	double L[3], iL[3];
	auto slot_of = [&](int i) {
		if(self >= 0 && i >= self) i++;
		int nb = 0;
		while(nb < 26 && i >= cn[nb]) {
			i -= cn[nb];
			nb++;
		}
		return st[nb] + i;
	};
	auto pos = [&](int i, double q[3]) {
		const int s2 = slot_of(i);
		for(int k = 0; k < 3; k++) q[k] = F[(size_t) k * S + s2];
	};
	auto ori = [&](int i, double Oq[9]) {
		const int s2 = slot_of(i);
		for(int k = 0; k < 9; k++) Oq[k] = F[(size_t) (3 + k) * S + s2];
	};
	if(b < m_pad && !live) {
		if(lane < NN_IN_PAD / 2) {
			row[2 * lane] = (__bf16) 0.f;
			row[2 * lane + 1] = (__bf16) 0.f;
		}
		if(lane < 6) target[(size_t) b * 6 + lane] = 0.f;
	}
	if(live) {
		const uint32_t k0 = (uint32_t) seed, k1 = (uint32_t) (seed >> 32);
		const u32x4 r0 = philox(u32x4{(uint32_t) b, 0xA5A50000u, (uint32_t) step, (uint32_t) (step >> 32)}, k0, k1);
		const u32x4 r1 = philox(u32x4{(uint32_t) b, 0xA5A50001u, (uint32_t) step, (uint32_t) (step >> 32)}, k0, k1);
		// a uniform slot of the whole context: a uniform particle of a replica
		// chosen with probability proportional to its size
		j = (int) __umulhi(r0.x, (uint32_t) d.Ntot);
		const Rep &rp = d.reps[d.prep[j]];
		F = pick_buf(d.f, rp.sel);
		for(int k = 0; k < 3; k++) {
			L[k] = rp.g.box[k];
			iL[k] = rp.g.inv_box[k];
		}
		double p[3], O[9];
		for(int k = 0; k < 3; k++) p[k] = F[(size_t) k * S + j];
		for(int k = 0; k < 9; k++) O[k] = F[(size_t) (3 + k) * S + j];
		// its cell on the current grid (particles never leave their cells
		// between rebuilds) and the 27 cells around it
		const int n0 = rp.g.n[0], n1 = rp.g.n[1], n2 = rp.g.n[2];
		const int ix = cell_coord(p[0], rp.g.shift[0], rp.g.inv_box[0], n0);
		const int iy = cell_coord(p[1], rp.g.shift[1], rp.g.inv_box[1], n1);
		const int iz = cell_coord(p[2], rp.g.shift[2], rp.g.inv_box[2], n2);
		for(int nb = 0; nb < 27; nb++) {
			const int cx = wrapi(ix + nb / 9 - 1, n0), cy = wrapi(iy + (nb / 3) % 3 - 1, n1),
			          cz = wrapi(iz + nb % 3 - 1, n2);
			const int cc = rp.cell_off + (cx * n1 + cy) * n2 + cz;
			st[nb] = d.start[cc];
			cn[nb] = d.start[cc + 1] - st[nb];
			if(self < 0 && j >= st[nb] && j < st[nb] + cn[nb]) self = tot + (j - st[nb]);
			tot += cn[nb];
		}
		// the perturbed state: Gaussian displacement, rotation by an angle
		// uniform in [0, theta_max) about a uniform axis
		double z[6];
		gauss2(r0.y, r0.z, z[0], z[1]);
		gauss2(r0.w, r1.x, z[2], z[3]);
		gauss2(r1.y, r1.z, z[4], z[5]);
		for(int k = 0; k < 3; k++) pt[k] = p[k] + pos_noise * z[k];
		double ax[3] = {z[3], z[4], z[5]};
		const double an = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
		for(int k = 0; k < 3; k++) ax[k] /= an;
		const double ang = u01(r1.w) * theta_max;
		double sn, cs, R[9];
		sincos(ang, &sn, &cs);
		rotation_matrix(ax, sn, cs, R);
		mat_mul3(O, R, Ot);
		if(lane == 0) {
			float *t = target + (size_t) b * 6;
			for(int k = 0; k < 3; k++)
				t[k] = (float) (Ot[3 * k] * (p[0] - pt[0]) + Ot[3 * k + 1] * (p[1] - pt[1]) + Ot[3 * k + 2] * (p[2] - pt[2]));
			for(int k = 0; k < 3; k++) t[3 + k] = (float) -ax[k];
		}
		nv = desc_collect<64>(pt, self >= 0 ? tot - 1 : tot, pos, L, iL, Ds[wave], row, lane);
	}
	__syncthreads();
	if(live) desc_fill<64>(pt, Ot, nv, pos, ori, L, iL, Ds[wave], row, lane);
}

void launch_train_samples(const DevState &d, uint64_t seed, uint64_t step, int m, double pos_noise, double theta_max,
                          __bf16 *x, float *target, hipStream_t s) {
	const int m_pad = (m + 31) & ~31;
	hipLaunchKernelGGL(k_train_samples, dim3((m_pad + TS_WAVES - 1) / TS_WAVES), dim3(64 * TS_WAVES), 0, s, d, seed,
	                   step, m, m_pad, pos_noise, theta_max, x, target);
}

}  // namespace ppd
