// filter_kernel.hip -- the FDK cosine pre-weight and ramp filter of
// feldkamp.m:103-132 (feldkamp_weight1) and fdk_filter.m:1-63, fused into one
// gfx950 kernel: weight, zero-padded FFT along s, multiplication by the
// filter's response, inverse FFT, crop, and the transpose to the [nt ns na]
// layout the backprojector takes.
//
// A workgroup of four waves takes eight detector rows of one view; each wave
// packs two rows as the real and imaginary parts of one complex signal of
// npad points in LDS.  The response H is real and even, so filtering the
// packed signal filters both rows: ifft(H fft(a + i b)) = h*a + i h*b.  The
// inverse transform is conj(fft(conj(.))) / npad, so both transforms are the
// same forward Stockham FFT (fft_core.h), in place: every pass reads its
// butterflies' inputs into registers, then writes them back.  In LDS every
// eight points are followed by one unused point, so that the strided writes
// of the early passes spread over the banks.  The twiddle table (each
// butterfly's twiddles contiguous) is built in FP64 on the host and kept per
// device and length, and stays in the caches.  HBM traffic is one read of the
// views and one write of the result (the weight table and the response stay
// in L2).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>
#include <utility>
#include <vector>

#include "fdkhip.h"
#include "fft_core.h"

namespace fdk {

constexpr int FLT_ROWS = 8;        // detector rows per workgroup (4 waves x 2)
constexpr int FLT_MAX_N = 4096;    // longest padded row

struct FilterArgs {
	const float *proj;       // [na][nt][ns] (s fastest)
	const float *ww1;        // [nt][ns] cosine pre-weight
	const float *resp;       // [npad] filter response (times ds)
	const fft::cf *tw;       // twiddles, pass by pass (fft::tw_offset)
	float *tsa;              // [na][ns][nt] (t fastest)
	int ns, nt, na, groups;  // groups: row groups per view
	int pairs;               // nt even and tsa 8-byte aligned: 8-byte stores
	float inv_n;             // 1 / npad
};

extern __shared__ fft::cf flt_lds[];

// LDS position of point i of a signal, and the points a signal takes: one
// unused point after every eight spreads the early passes' strided writes
// over the banks
__device__ __forceinline__ int flt_pos(int i) { return i + (i >> 3); }
#define FLT_PADDED(n) ((n) + (n) / 8)

// Orders one wave's LDS accesses: a wave's LDS instructions execute in
// program order, so a wave that owns its signal needs no workgroup barrier
// between passes, only this fence against compiler reordering.  The fence
// covers LDS only, so global loads issued ahead (the next pass's twiddles)
// stay in flight across it.
__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// The shape of the pass of span NS over N points, as one wave runs it
template <int N, int NS>
struct PassShape {
	static constexpr int R = fft::pass_radix(N, NS);
	static constexpr int NB = N / R;                 // butterflies
	static constexpr int PER = (NB + 63) / 64;       // per lane
	static constexpr bool LAST = NS * R == N;
	// butterfly b of a lane (lanes past the last butterfly repeat it: their
	// loads stay in bounds and their stores are skipped)
	static __device__ __forceinline__ int bfly(int lane, int b)
	{
		const int j = lane + 64 * b;
		return (NB % 64 == 0 || j < NB) ? j : NB - 1;
	}
	static __device__ __forceinline__ bool owns(int lane, int b)
	{
		return NB % 64 == 0 || lane + 64 * b < NB;
	}
};

// The twiddles of a lane's butterflies in the pass of span NS (none for the
// first pass), loaded a pass ahead of their use: each butterfly's R twiddles
// are contiguous in the table, so they arrive as a few 16-byte loads.
template <int N, int NS, bool ANY = (NS > 1 && NS < N)>
struct PassTwiddles {
	__device__ __forceinline__ void load(const fft::cf *, int) {}
	__device__ __forceinline__ const fft::cf *at(int) const { return nullptr; }
};

template <int N, int NS>
struct PassTwiddles<N, NS, true> {
	using P = PassShape<N, NS>;
	fft::cf w[P::PER][P::R];
	__device__ __forceinline__ void load(const fft::cf *tw, int lane)
	{
		constexpr int TW0 = fft::tw_offset(N, NS);
#pragma unroll
		for (int b = 0; b < P::PER; ++b) {
			const fft::cf *p = tw + TW0 + (P::bfly(lane, b) % NS) * P::R;
#pragma unroll
			for (int m = 0; m < P::R; ++m)
				w[b][m] = p[m];
		}
	}
	__device__ __forceinline__ const fft::cf *at(int b) const { return w[b]; }
};

// the span of the last pass of an N-point transform
constexpr int last_span(int N)
{
	int ns = 1;
	while (ns * fft::pass_radix(N, ns) < N)
		ns *= fft::pass_radix(N, ns);
	return ns;
}

// Per-output factors of a transform's last pass, loaded a pass ahead like
// the twiddles: none for the inverse transform; the filter response H[k] / N
// for the forward one
struct NoOutputScale {
	__device__ __forceinline__ void load(int) {}
	__device__ __forceinline__ float at(int, int) const { return 1.f; }
};

template <int N>
struct ResponseScale {
	static constexpr int NSL = last_span(N);
	using P = PassShape<N, NSL>;
	const float *resp;
	float inv_n;
	float h[P::PER][P::R];
	// output m of butterfly j of the last pass is point j + m * NSL
	__device__ __forceinline__ void load(int lane)
	{
#pragma unroll
		for (int b = 0; b < P::PER; ++b) {
			const int j = P::bfly(lane, b);
#pragma unroll
			for (int m = 0; m < P::R; ++m)
				h[b][m] = resp[j + m * NSL];
		}
	}
	// scaled where it is used, so that the loads stay in flight until then
	__device__ __forceinline__ float at(int b, int m) const { return h[b][m] * inv_n; }
};

// One radix-R pass of span NS over N points: butterfly j's inputs come from
// load(index), its outputs go to store(index, value, scale).  IN_LDS: the
// inputs are read from the buffer the outputs are written to (in place), so
// every read of the pass precedes the first write.  HALF_IN: inputs m >= R/2
// are zero and not loaded; HALF_OUT: outputs m >= R/2 are not needed (the
// first pass over, and the last pass back to, a row at most N/2 long).  The
// signal is the wave's own.
template <int N, int NS, bool IN_LDS, bool HALF_IN, bool HALF_OUT, class Load, class Store,
	class Scale>
__device__ __forceinline__ void fft_pass(const PassTwiddles<N, NS> &tw, const Scale &scale, int lane,
	Load load, Store store)
{
	using P = PassShape<N, NS>;
	constexpr int R = P::R;
	constexpr int NIN = HALF_IN ? R / 2 : R, NOUT = HALF_OUT ? R / 2 : R;
	fft::cf v[P::PER][R];
#pragma unroll
	for (int b = 0; b < P::PER; ++b) {
		const int j = P::bfly(lane, b);
#pragma unroll
		for (int m = 0; m < R; ++m)
			v[b][m] = m < NIN ? load(fft::in_index(j, m, N, R)) : fft::cf{0.f, 0.f};
	}
	if constexpr (IN_LDS)
		wave_sync();
#pragma unroll
	for (int b = 0; b < P::PER; ++b) {
		const int j = P::bfly(lane, b);
		fft::butterfly<R, HALF_IN>(v[b], tw.at(b));
		if (P::owns(lane, b)) {
#pragma unroll
			for (int m = 0; m < NOUT; ++m)
				store(fft::out_index(j, m, NS, R), v[b][m], scale.at(b, m));
		}
	}
	wave_sync();
}
// Forward DFT of N points (natural order in and out) over the signal at buf:
// the first pass takes its inputs from first_load, the last pass gives its
// outputs to last_store with scale's factors, the passes between work in
// place in buf.  Each pass's twiddles (and the last pass's factors) are
// loaded while the pass before it runs.
template <int N, int NS, bool FIRST, bool HALF_IN, bool HALF_OUT, class Load, class Store,
	class Scale>
__device__ __forceinline__ void fft_forward(fft::cf *buf, const fft::cf *tw, int lane,
	Load first_load, Store last_store, Scale &scale, const PassTwiddles<N, NS> &wcur)
{
	if constexpr (NS < N) {
		using P = PassShape<N, NS>;
		constexpr int NEXT = NS * P::R;
		PassTwiddles<N, NEXT> wnext;
		wnext.load(tw, lane);
		if constexpr (FIRST && P::LAST)
			scale.load(lane);
		if constexpr (!P::LAST && PassShape<N, NEXT>::LAST)
			scale.load(lane);
		constexpr bool HI = FIRST && HALF_IN, HO = P::LAST && HALF_OUT;
		auto lds_load = [buf](int i) { return buf[flt_pos(i)]; };
		auto lds_store = [buf](int i, fft::cf z, float) { buf[flt_pos(i)] = z; };
		const NoOutputScale none;
		if constexpr (FIRST && P::LAST)
			fft_pass<N, NS, false, HI, HO>(wcur, scale, lane, first_load, last_store);
		else if constexpr (FIRST)
			fft_pass<N, NS, false, HI, HO>(wcur, none, lane, first_load, lds_store);
		else if constexpr (P::LAST)
			fft_pass<N, NS, true, HI, HO>(wcur, scale, lane, lds_load, last_store);
		else
			fft_pass<N, NS, true, HI, HO>(wcur, none, lane, lds_load, lds_store);
		fft_forward<N, NEXT, false, HALF_IN, HALF_OUT>(buf, tw, lane, first_load, last_store,
			scale, wnext);
	}
}

// Waves per SIMD that the LDS of the four signals leaves room for: more
// registers do not buy occupancy past it, so the scheduler may spend them on
// keeping a pass's loads in flight
constexpr int flt_waves(int N)
{
	const int blocks = (160 * 1024) / (4 * (N + N / 8) * 8);
	return blocks < 1 ? 1 : (blocks > 8 ? 8 : blocks);
}

// HALF: ns <= N/2, so the upper half of every padded row is zero and the
// upper half of every filtered row is cropped away.
template <int N, bool HALF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, flt_waves(N))))
filter_kernel(const FilterArgs a)
{
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
	const int v = blockIdx.x / a.groups;
	const int t0 = (blockIdx.x - v * a.groups) * FLT_ROWS;
	fft::cf *buf = flt_lds + wave * FLT_PADDED(N);
	const int ta = t0 + 2 * wave, tb = ta + 1;
	const bool ha = ta < a.nt, hb = tb < a.nt;
	// rows past nt read row 0 of the view and are zeroed
	const float *pa = a.proj + ((size_t) v * a.nt + (ha ? ta : 0)) * a.ns;
	const float *pb = a.proj + ((size_t) v * a.nt + (hb ? tb : 0)) * a.ns;
	const float *wa = a.ww1 + (size_t) (ha ? ta : 0) * a.ns;
	const float *wb = a.ww1 + (size_t) (hb ? tb : 0) * a.ns;
	const int ns = a.ns;
	// the weighted rows, zero-padded to N, straight into the first pass
	// (branch-free: past the row end the load is clamped and the value zeroed)
	auto weighted = [=](int s) {
		const bool in = s < ns;
		const int sc = in ? s : ns - 1;
		const float xa = pa[sc] * wa[sc], xb = pb[sc] * wb[sc];
		return fft::cf{(in & ha) ? xa : 0.f, (in & hb) ? xb : 0.f};
	};
	// the forward transform's last pass multiplies by H / N and conjugates:
	// the second forward transform is then the conjugated inverse
	ResponseScale<N> hs;
	hs.resp = a.resp;
	hs.inv_n = a.inv_n;
	auto filtered = [buf](int k, fft::cf z, float h) {
		buf[flt_pos(k)] = fft::cf{z.x * h, -z.y * h};
	};
	const PassTwiddles<N, 1> w0;
	fft_forward<N, 1, true, HALF, false>(buf, a.tw, lane, weighted, filtered, hs, w0);
	// inverse (outputs past ns are written, never read)
	auto lds_load = [buf](int i) { return buf[flt_pos(i)]; };
	auto lds_store = [buf](int i, fft::cf z, float) { buf[flt_pos(i)] = z; };
	NoOutputScale none;
	fft_forward<N, 1, false, false, HALF>(buf, a.tw, lane, lds_load, lds_store, none, w0);
	__syncthreads();   // the transpose reads all four waves' signals
	// row ta is the real part, row tb minus the imaginary part; out [nt ns na]
	// with the group's eight rows contiguous per s
	float *out = a.tsa + (size_t) v * ns * a.nt;
	if (a.pairs) {
		// even nt: each lane moves one signal's pair of rows as one 8-byte word
#pragma unroll 4
		for (int i = threadIdx.x; i < ns * (FLT_ROWS / 2); i += blockDim.x) {
			const int s = i >> 2, p = i & 3, t = t0 + 2 * p;
			if (t < a.nt) {
				const fft::cf z = flt_lds[p * FLT_PADDED(N) + flt_pos(s)];
				*reinterpret_cast<float2 *>(out + (size_t) s * a.nt + t) = float2(z.x, -z.y);
			}
		}
	} else {
#pragma unroll 4
		for (int i = threadIdx.x; i < ns * FLT_ROWS; i += blockDim.x) {
			const int s = i >> 3, tt = i & 7, t = t0 + tt;
			if (t < a.nt) {
				const fft::cf z = flt_lds[(tt >> 1) * FLT_PADDED(N) + flt_pos(s)];
				out[(size_t) s * a.nt + t] = (tt & 1) ? -z.y : z.x;
			}
		}
	}
}

template <int N>
static hipError_t launch_filter_n(const FilterArgs &a, hipStream_t s)
{
	const size_t lds = (size_t) 4 * FLT_PADDED(N) * sizeof(fft::cf);   // four signals
	const dim3 grid((unsigned) (a.na * a.groups));
	if (2 * a.ns <= N)
		hipLaunchKernelGGL((filter_kernel<N, true>), grid, dim3(256), lds, s, a);
	else
		hipLaunchKernelGGL((filter_kernel<N, false>), grid, dim3(256), lds, s, a);
	return hipGetLastError();
}

static hipError_t launch_filter(const FilterArgs &a, int n, hipStream_t s)
{
	switch (n) {
	case 2: return launch_filter_n<2>(a, s);
	case 4: return launch_filter_n<4>(a, s);
	case 8: return launch_filter_n<8>(a, s);
	case 16: return launch_filter_n<16>(a, s);
	case 32: return launch_filter_n<32>(a, s);
	case 64: return launch_filter_n<64>(a, s);
	case 128: return launch_filter_n<128>(a, s);
	case 256: return launch_filter_n<256>(a, s);
	case 512: return launch_filter_n<512>(a, s);
	case 1024: return launch_filter_n<1024>(a, s);
	case 2048: return launch_filter_n<2048>(a, s);
	case 4096: return launch_filter_n<4096>(a, s);
	default: return hipErrorInvalidValue;
	}
}

// the twiddle table of length n on the current device, built on first use
// and kept for the life of the process
static std::mutex g_tw_mu;
static std::vector<std::pair<long, fft::cf *>> g_tw;   // (device * 2^20 + n, table)

static hipError_t twiddles(int n, const fft::cf **out)
{
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess)
		return e;
	const long key = (long) dev * (1l << 20) + n;
	std::lock_guard<std::mutex> lk(g_tw_mu);
	for (auto &kv : g_tw)
		if (kv.first == key) {
			*out = kv.second;
			return hipSuccess;
		}
	// the pass-by-pass table of fft_core.h, each entry rounded from FP64
	const int ne = fft::tw_entries(n);
	std::vector<fft::cf> h((size_t) (ne > 0 ? ne : 1));
	for (int ns = 1; ns < n; ns *= fft::pass_radix(n, ns)) {
		const int r = fft::pass_radix(n, ns), off = fft::tw_offset(n, ns);
		for (int j = 0; j < ns; ++j)
			for (int m = 0; m < r; ++m) {
				const double ang = -2.0 * M_PI * (double) m * j / ((double) ns * r);
				h[(size_t) off + j * r + m] = fft::cf{(float) cos(ang), (float) sin(ang)};
			}
	}
	fft::cf *d = nullptr;
	if ((e = hipMalloc(&d, sizeof(fft::cf) * h.size())) != hipSuccess)
		return e;
	if ((e = hipMemcpy(d, h.data(), sizeof(fft::cf) * h.size(), hipMemcpyHostToDevice)) != hipSuccess) {
		(void) hipFree(d);
		return e;
	}
	g_tw.emplace_back(key, d);
	*out = d;
	return hipSuccess;
}

} // namespace fdk

using namespace fdk;

int fdkhip_fail_msg(int code, const char *msg);   // fdk_api.hip

extern "C" int fdkhip_fdk_filter_views(const float *proj, int32_t ns, int32_t nt, int32_t na,
	const float *ww1, const float *response, int32_t npad, float *tsa, void *stream)
{
	if (ns < 0 || nt < 0 || na < 0)
		return fdkhip_fail_msg(FDKHIP_EARG, "negative size");
	if (npad < 2 || npad > FLT_MAX_N || (npad & (npad - 1)))
		return fdkhip_fail_msg(npad > FLT_MAX_N ? FDKHIP_EUNSUPPORTED : FDKHIP_EARG,
			"npad must be a power of two from 2 to 4096");
	if (npad < ns)
		return fdkhip_fail_msg(FDKHIP_EARG, "npad must be at least ns");
	if (na == 0 || (long) ns * nt == 0)
		return FDKHIP_OK;
	if (!proj || !ww1 || !response || !tsa)
		return fdkhip_fail_msg(FDKHIP_EARG, "NULL device pointer");
	if (proj == tsa)
		return fdkhip_fail_msg(FDKHIP_EARG, "tsa must not alias proj");
	if ((long) na * ((nt + FLT_ROWS - 1) / FLT_ROWS) > 0x7fffffffL)
		return fdkhip_fail_msg(FDKHIP_EARG, "too many views");
	const fft::cf *tw = nullptr;
	hipError_t e = twiddles(npad, &tw);
	if (e != hipSuccess)
		return fdkhip_fail_msg(FDKHIP_EHIP, hipGetErrorString(e));
	FilterArgs a;
	a.proj = proj;
	a.ww1 = ww1;
	a.resp = response;
	a.tw = tw;
	a.tsa = tsa;
	a.ns = ns;
	a.nt = nt;
	a.na = na;
	a.groups = (nt + FLT_ROWS - 1) / FLT_ROWS;
	a.pairs = (nt % 2 == 0) && ((uintptr_t) tsa % 8 == 0);
	a.inv_n = 1.f / (float) npad;
	e = launch_filter(a, npad, (hipStream_t) stream);
	if (e != hipSuccess)
		return fdkhip_fail_msg(FDKHIP_EHIP, hipGetErrorString(e));
	return FDKHIP_OK;
}
