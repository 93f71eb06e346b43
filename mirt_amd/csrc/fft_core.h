// fft_core.h -- the complex FFT passes of the fused FDK filter kernel
// (filter_kernel.hip), written as host/device functions so that the
// repository's CPU tests run the very same index and butterfly arithmetic
// against numpy's FFT.
//
// Stockham autosort FFT (the formulation of Govindaraju, Lloyd, Dotsenko,
// Smith and Manferdelli, "High performance discrete Fourier transforms on
// graphics processors", SC 2008): for a length N = R^k * r, passes of radix
// R with span Ns = 1, R, R^2, ...; butterfly j of a pass reads the R
// elements j + m*N/R, multiplies element m by w^(m * (j mod Ns)) with
// w = exp(-2 pi i / (Ns R)), takes an R-point DFT and writes element m to
// (j / Ns) * Ns * R + (j mod Ns) + m * Ns.  The result is the DFT in natural
// order.  Twiddles come from a table laid out pass by pass (tw_offset).
#pragma once

#if defined(__HIPCC__)
#define FFT_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define FFT_HD inline
#endif

namespace fft {

struct alignas(8) cf {
	float x, y;
};

FFT_HD cf add(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
FFT_HD cf sub(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
// complex product with fused multiply-adds (two products, two FMAs)
FFT_HD cf mul(cf a, cf b)
{
	return cf{fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x)};
}
// a * (-i)
FFT_HD cf mul_mi(cf a) { return cf{a.y, -a.x}; }

// in-place 2-, 4- and 8-point DFTs (forward, exp(-2 pi i ...)), natural order
FFT_HD void dft2(cf &a, cf &b)
{
	const cf t = a;
	a = add(t, b);
	b = sub(t, b);
}

FFT_HD void dft4(cf &v0, cf &v1, cf &v2, cf &v3)
{
	dft2(v0, v2);                 // v0 = x0 + x2, v2 = x0 - x2
	dft2(v1, v3);                 // v1 = x1 + x3, v3 = x1 - x3
	v3 = mul_mi(v3);              // (x1 - x3) * -i
	dft2(v0, v1);                 // X0, X2
	dft2(v2, v3);                 // X1, X3
	const cf x1 = v2, x2 = v1;
	v1 = x1;
	v2 = x2;
}

// the 4-point DFT of (x0, x1, 0, 0)
FFT_HD void dft4_half(cf &v0, cf &v1, cf &v2, cf &v3)
{
	const cf x0 = v0, x1 = v1;
	v0 = add(x0, x1);
	v2 = sub(x0, x1);
	v1 = add(x0, mul_mi(x1));
	v3 = sub(x0, mul_mi(x1));
}

template <bool HALF = false>
FFT_HD void dft8(cf (&v)[8])
{
	const float r = 0.70710678118654752440f;   // sqrt(2)/2
	// split into even and odd samples: two 4-point DFTs, then combine
	cf e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
	cf o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
	if (HALF) {                   // v[4..7] are zero
		dft4_half(e0, e1, e2, e3);
		dft4_half(o0, o1, o2, o3);
	} else {
		dft4(e0, e1, e2, e3);
		dft4(o0, o1, o2, o3);
	}
	// o_k * exp(-2 pi i k / 8)
	o1 = cf{r * (o1.x + o1.y), r * (o1.y - o1.x)};
	o2 = mul_mi(o2);
	o3 = cf{r * (o3.y - o3.x), -r * (o3.x + o3.y)};
	v[0] = add(e0, o0);
	v[4] = sub(e0, o0);
	v[1] = add(e1, o1);
	v[5] = sub(e1, o1);
	v[2] = add(e2, o2);
	v[6] = sub(e2, o2);
	v[3] = add(e3, o3);
	v[7] = sub(e3, o3);
}

// element indices of butterfly j of a radix-R pass with span Ns over N points
FFT_HD int in_index(int j, int m, int N, int R) { return j + m * (N / R); }
FFT_HD int out_index(int j, int m, int Ns, int R) { return (j / Ns) * Ns * R + (j % Ns) + m * Ns; }

// the radix of each pass for N = 2^k: radix-8 passes, then one pass of 2 or 4
FFT_HD constexpr int pass_radix(int N, int Ns) { return (N / Ns) % 8 == 0 ? 8 : N / Ns; }

// Twiddles, pass by pass: the pass of span Ns and radix R owns Ns * R entries
// at offset tw_offset(N, Ns); entry (j mod Ns) * R + m is
// exp(-2 pi i m (j mod Ns) / (Ns R)), so a butterfly's R twiddles are
// contiguous.  The passes of span 1 have no twiddles (offset unused).
FFT_HD constexpr int tw_offset(int N, int Ns)
{
	int off = 0;
	for (int s = 1; s < Ns; s *= pass_radix(N, s))
		off += s * pass_radix(N, s);
	return off;
}
FFT_HD constexpr int tw_entries(int N) { return tw_offset(N, N); }

// one butterfly of a radix-R pass, in registers: twiddles w[1..R-1] (none
// when w is null), then the R-point DFT.  HALF: the inputs v[R/2..R-1] are
// zero (the first pass over a signal zero-padded to twice its length or
// more), which the radix-8 DFT skips.
template <int R, bool HALF = false>
FFT_HD void butterfly(cf (&v)[R], const cf *w)
{
	if (w) {
		for (int m = 1; m < R; ++m)
			v[m] = mul(v[m], w[m]);
	}
	if constexpr (R == 8) {
		dft8<HALF>(v);
	} else if constexpr (R == 4) {
		dft4(v[0], v[1], v[2], v[3]);
	} else {
		dft2(v[0], v[1]);
	}
}

} // namespace fft
