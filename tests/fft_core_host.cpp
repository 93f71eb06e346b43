// TEST INFRASTRUCTURE: the FFT passes of mirt_amd/csrc/fft_core.h run on the
// host (the device kernel's index and butterfly arithmetic), for
// tests/test_host.py to compare with numpy's FFT.
#include <cmath>
#include <vector>
#include "fft_core.h"

// one pass; HALF: the inputs of index >= N/2 are zero and not read (the
// kernel's first pass over a row zero-padded to twice its length or more)
template <int R, bool HALF = false>
static void pass(const fft::cf *in, fft::cf *out, int N, int Ns, const fft::cf *tw)
{
	const fft::cf *tp = tw + fft::tw_offset(N, Ns);
	for (int j = 0; j < N / R; ++j) {
		fft::cf v[R];
		for (int m = 0; m < R; ++m)
			v[m] = (HALF && m >= R / 2) ? fft::cf{0.f, 0.f} : in[fft::in_index(j, m, N, R)];
		fft::butterfly<R, HALF>(v, Ns > 1 ? tp + (j % Ns) * R : nullptr);
		for (int m = 0; m < R; ++m)
			out[fft::out_index(j, m, Ns, R)] = v[m];
	}
}

// the pass-by-pass twiddle table of fft_core.h for length N, as the library
// builds it (FP64, rounded): returns the number of entries written to tw
extern "C" int fft_core_twiddles(int N, float *tw)
{
	auto *t = reinterpret_cast<fft::cf *>(tw);
	for (int Ns = 1; Ns < N; Ns *= fft::pass_radix(N, Ns)) {
		const int R = fft::pass_radix(N, Ns), off = fft::tw_offset(N, Ns);
		for (int r = 0; r < Ns; ++r)
			for (int m = 0; m < R; ++m) {
				const double a = -2.0 * M_PI * (double) m * r / ((double) Ns * R);
				t[off + r * R + m] = fft::cf{(float) std::cos(a), (float) std::sin(a)};
			}
	}
	return fft::tw_entries(N);
}

// forward DFT of x[N] (interleaved re, im), N a power of two >= 2, with the
// twiddle table of fft_core_twiddles; the result replaces x.  half != 0: the
// first pass reads only x[0 .. N/2) and takes the rest as zero.
extern "C" void fft_core_forward(float *x, int N, const float *tw, int half)
{
	auto *a = reinterpret_cast<fft::cf *>(x);
	const auto *t = reinterpret_cast<const fft::cf *>(tw);
	std::vector<fft::cf> b(N);
	fft::cf *src = a, *dst = b.data();
	for (int Ns = 1; Ns < N;) {
		const int R = fft::pass_radix(N, Ns);
		const bool h = half && Ns == 1;
		if (R == 8)
			h ? pass<8, true>(src, dst, N, Ns, t) : pass<8>(src, dst, N, Ns, t);
		else if (R == 4)
			h ? pass<4, true>(src, dst, N, Ns, t) : pass<4>(src, dst, N, Ns, t);
		else
			h ? pass<2, true>(src, dst, N, Ns, t) : pass<2>(src, dst, N, Ns, t);
		Ns *= R;
		fft::cf *tmp = src;
		src = dst;
		dst = tmp;
	}
	if (src != a)
		for (int i = 0; i < N; ++i)
			a[i] = src[i];
}
