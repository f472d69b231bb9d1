// LDS-resident Stockham FFT building blocks for power-of-two lengths (gfx950).
//
// A workgroup holds NTR transforms of length N in LDS as s[row * LD + col]
// (row = index along the transform, col = transform id, LD >= NTR odd so that
// 16-byte complex128 lanes spread over the LDS banks).  Each Stockham pass
// reads RAD points per butterfly into registers, applies the twiddles and a
// radix-RAD DFT, and writes them back in autosort order, with a barrier
// between the read and the write phase.  Radix 8 is used first, then 4 or 2.
//
// Twiddles come from a table tw[m] = exp(-2 pi i m / kTwMax) built on the host
// in long double; forward transforms are unscaled, inverse transforms use the
// conjugate twiddles and are also unscaled (the 1/P of the inverse is folded
// into the spectral G0 table).
#pragma once

#include "msc_common.hpp"

namespace msc {

constexpr int kTwMax = 1024;  // largest supported transform length

template <int NTR>
struct lds_ld {
  static constexpr int value = (NTR % 2 == 0) ? NTR + 1 : NTR;
};

template <typename R, bool INV>
__device__ __forceinline__ cx<R> mul_mi(cx<R> a) {
  // forward: a * (-i); inverse: a * (+i)
  return INV ? mk<R>(-a.y, a.x) : mk<R>(a.y, -a.x);
}

template <typename R, bool INV>
__device__ __forceinline__ void dft2(cx<R>& a, cx<R>& b) {
  cx<R> t = a;
  a = cadd<R>(t, b);
  b = csub<R>(t, b);
}

template <typename R, bool INV>
__device__ __forceinline__ void dft4(cx<R>& x0, cx<R>& x1, cx<R>& x2, cx<R>& x3) {
  cx<R> t0 = cadd<R>(x0, x2), t1 = csub<R>(x0, x2);
  cx<R> t2 = cadd<R>(x1, x3), t3 = mul_mi<R, INV>(csub<R>(x1, x3));
  x0 = cadd<R>(t0, t2);
  x2 = csub<R>(t0, t2);
  x1 = cadd<R>(t1, t3);
  x3 = csub<R>(t1, t3);
}

template <typename R, bool INV>
__device__ __forceinline__ void dft8(cx<R>* v) {
  // even/odd split: E = DFT4(v0,v2,v4,v6), O = DFT4(v1,v3,v5,v7)
  cx<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cx<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<R, INV>(e0, e1, e2, e3);
  dft4<R, INV>(o0, o1, o2, o3);
  const R h = R(0.70710678118654752440084436210485);
  // W8^1 = (1 -+ i)/sqrt2, W8^2 = -+i, W8^3 = (-1 -+ i)/sqrt2
  cx<R> w1 = INV ? mk<R>(h * (o1.x - o1.y), h * (o1.x + o1.y))
                 : mk<R>(h * (o1.x + o1.y), h * (o1.y - o1.x));
  cx<R> w2 = mul_mi<R, INV>(o2);
  cx<R> w3 = INV ? mk<R>(-h * (o3.x + o3.y), h * (o3.x - o3.y))
                 : mk<R>(h * (o3.y - o3.x), -h * (o3.x + o3.y));
  v[0] = cadd<R>(e0, o0);
  v[4] = csub<R>(e0, o0);
  v[1] = cadd<R>(e1, w1);
  v[5] = csub<R>(e1, w1);
  v[2] = cadd<R>(e2, w2);
  v[6] = csub<R>(e2, w2);
  v[3] = cadd<R>(e3, w3);
  v[7] = csub<R>(e3, w3);
}

template <typename R, int RAD, bool INV>
__device__ __forceinline__ void dft(cx<R>* v) {
  if constexpr (RAD == 2) {
    dft2<R, INV>(v[0], v[1]);
  } else if constexpr (RAD == 4) {
    dft4<R, INV>(v[0], v[1], v[2], v[3]);
  } else {
    static_assert(RAD == 8, "radix must be 2, 4 or 8");
    dft8<R, INV>(v);
  }
}

// dft() of inputs whose upper half v[RAD/2 ..] is zero (not read): the zero
// terms are left out instead of added (x + 0 is not x for x = -0, so the
// compiler keeps such additions).
template <typename R, bool INV>
__device__ __forceinline__ void dft4_half(cx<R>& x0, cx<R>& x1, cx<R>& x2, cx<R>& x3) {
  const cx<R> a = x0, m = mul_mi<R, INV>(x1);
  x2 = csub<R>(a, x1);
  x0 = cadd<R>(a, x1);
  x1 = cadd<R>(a, m);
  x3 = csub<R>(a, m);
}

template <typename R, int RAD, bool INV>
__device__ __forceinline__ void dft_half(cx<R>* v) {
  if constexpr (RAD == 2) {
    v[1] = v[0];
  } else if constexpr (RAD == 4) {
    dft4_half<R, INV>(v[0], v[1], v[2], v[3]);
  } else {
    static_assert(RAD == 8, "radix must be 2, 4 or 8");
    cx<R> e0 = v[0], e1 = v[2], e2, e3;
    cx<R> o0 = v[1], o1 = v[3], o2, o3;
    dft4_half<R, INV>(e0, e1, e2, e3);
    dft4_half<R, INV>(o0, o1, o2, o3);
    const R h = R(0.70710678118654752440084436210485);
    cx<R> w1 = INV ? mk<R>(h * (o1.x - o1.y), h * (o1.x + o1.y))
                   : mk<R>(h * (o1.x + o1.y), h * (o1.y - o1.x));
    cx<R> w2 = mul_mi<R, INV>(o2);
    cx<R> w3 = INV ? mk<R>(-h * (o3.x + o3.y), h * (o3.x - o3.y))
                   : mk<R>(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    v[0] = cadd<R>(e0, o0);
    v[4] = csub<R>(e0, o0);
    v[1] = cadd<R>(e1, w1);
    v[5] = csub<R>(e1, w1);
    v[2] = cadd<R>(e2, w2);
    v[6] = csub<R>(e2, w2);
    v[3] = cadd<R>(e3, w3);
    v[7] = csub<R>(e3, w3);
  }
}

// One Stockham pass: sub-transform length NS grows to NS*RAD.
template <typename R, int N, int RAD, int NS, bool INV, int NTR, int LD, int NT>
__device__ __forceinline__ void stockham_pass(cx<R>* s, const cx<R>* __restrict__ tw) {
  constexpr int NB = N / RAD;
  constexpr int TOT = NB * NTR;
  constexpr int ITEMS = (TOT + NT - 1) / NT;
  constexpr int TWS = kTwMax / (NS * RAD);
  cx<R> v[ITEMS][RAD];
  const int tid = threadIdx.x;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = tid + it * NT;
    if (ITEMS * NT == TOT || w < TOT) {
      const int col = w % NTR;
      const int j = w / NTR;
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[it][r] = s[(j + r * NB) * LD + col];
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          cx<R> t = tw[r * k * TWS];
          if (INV) t.y = -t.y;
          v[it][r] = cmul<R>(v[it][r], t);
        }
      }
      dft<R, RAD, INV>(v[it]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = tid + it * NT;
    if (ITEMS * NT == TOT || w < TOT) {
      const int col = w % NTR;
      const int j = w / NTR;
      const int k = j % NS;
      const int base = (j / NS) * NS * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) s[(base + r * NS) * LD + col] = v[it][r];
    }
  }
  __syncthreads();
}

template <int REM>
struct pick_radix {
  static constexpr int value = (REM % 8 == 0) ? 8 : ((REM % 4 == 0) ? 4 : 2);
};

template <typename R, int N, int NS, bool INV, int NTR, int LD, int NT>
__device__ __forceinline__ void fft_stages(cx<R>* s, const cx<R>* __restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int RAD = pick_radix<N / NS>::value;
    stockham_pass<R, N, RAD, NS, INV, NTR, LD, NT>(s, tw);
    fft_stages<R, N, NS * RAD, INV, NTR, LD, NT>(s, tw);
  }
}

// In-LDS FFT of NTR columns of length N (forward if !INV).  Callers must have
// issued a barrier after filling s; the routine ends with a barrier.
template <typename R, int N, bool INV, int NTR, int LD, int NT>
__device__ __forceinline__ void lds_fft(cx<R>* s, const cx<R>* __restrict__ tw) {
  static_assert((N & (N - 1)) == 0 && N >= 2 && N <= kTwMax, "N must be a power of two");
  fft_stages<R, N, 1, INV, NTR, LD, NT>(s, tw);
}

// Sub-transform length NS entering the last Stockham pass of length N (the
// radix schedule of fft_stages).
constexpr int last_ns(int N) {
  int ns = 1;
  while (true) {
    const int rem = N / ns;
    const int rad = rem % 8 == 0 ? 8 : (rem % 4 == 0 ? 4 : 2);
    if (ns * rad == N) return ns;
    ns *= rad;
  }
}

// Every pass of fft_stages except the last one (ends with a barrier when it
// runs any pass).
template <typename R, int N, int NS, bool INV, int NTR, int LD, int NT>
__device__ __forceinline__ void fft_stages_but_last(cx<R>* s, const cx<R>* __restrict__ tw) {
  if constexpr (NS < last_ns(N)) {
    constexpr int RAD = pick_radix<N / NS>::value;
    stockham_pass<R, N, RAD, NS, INV, NTR, LD, NT>(s, tw);
    fft_stages_but_last<R, N, NS * RAD, INV, NTR, LD, NT>(s, tw);
  }
}

// First forward pass (NS = 1) of a transform whose input rows >= nrows are
// zero, nrows <= N / 2, read from src[row * NTR + col] (unpadded) and written
// to dst[row * LDD + col]: only the rows < nrows are read (rows r * N / RAD
// with r >= RAD / 2 are >= N / 2).  No barrier between read and write
// (src and dst are distinct); none at the end either.
template <typename R, int N, int NTR, int LDD, int NT>
__device__ __forceinline__ void first_pass_pruned(const cx<R>* src, cx<R>* dst, int nrows) {
  constexpr int RAD = pick_radix<N>::value;
  constexpr int NB = N / RAD;
  constexpr int TOT = NB * NTR;
  constexpr int ITEMS = (TOT + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = (int)threadIdx.x + it * NT;
    if (ITEMS * NT == TOT || w < TOT) {
      const int col = w % NTR;
      const int j = w / NTR;
      cx<R> v[RAD];
#pragma unroll
      for (int r = 0; r < RAD; ++r) {
        const int row = j + r * NB;
        v[r] = (r < RAD / 2 && row < nrows) ? src[row * NTR + col] : czero<R>();
      }
      dft_half<R, RAD, false>(v);
#pragma unroll
      for (int r = 0; r < RAD; ++r) dst[(j * RAD + r) * LDD + col] = v[r];
    }
  }
}

// The same first pass reading the rows < nrows straight from global memory,
// src[row * stride + col] (issued together before any is used), and writing
// the pass's output to s[row * LDD + col].  No barrier.
template <typename R, int N, int NTR, int LDD, int NT>
__device__ __forceinline__ void first_pass_pruned_global(const cx<R>* __restrict__ src,
                                                         size_t stride, cx<R>* dst, int nrows) {
  constexpr int RAD = pick_radix<N>::value;
  constexpr int NB = N / RAD;
  constexpr int TOT = NB * NTR;
  constexpr int ITEMS = (TOT + NT - 1) / NT;
  constexpr int HR = RAD > 1 ? RAD / 2 : 1;
  cx<R> v[ITEMS][RAD];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = (int)threadIdx.x + it * NT;
    const bool in = ITEMS * NT == TOT || w < TOT;
    const int col = w % NTR;
    const int j = w / NTR;
#pragma unroll
    for (int r = 0; r < RAD; ++r) {
      const int row = j + r * NB;
      const bool ok = in && r < HR && row < nrows;
      const cx<R> x = src[ok ? (size_t)row * stride + col : 0];
      v[it][r] = ok ? x : czero<R>();
    }
  }
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = (int)threadIdx.x + it * NT;
    if (ITEMS * NT == TOT || w < TOT) {
      const int col = w % NTR;
      const int j = w / NTR;
      if constexpr (RAD > 1) dft_half<R, RAD, false>(v[it]);
#pragma unroll
      for (int r = 0; r < RAD; ++r) dst[(j * RAD + r) * LDD + col] = v[it][r];
    }
  }
}

// Last pass of an inverse transform held in s (after fft_stages_but_last),
// handing only the output rows < nrows to store(row, col, value) instead of
// writing them back to LDS.  nrows <= N / 2: the outputs j + r NS with
// r >= RAD / 2 (rows >= N / 2) are never formed.
template <typename R, int N, int NTR, int LD, int NT, typename Store>
__device__ __forceinline__ void last_pass_inv_store(const cx<R>* s, const cx<R>* __restrict__ tw,
                                                    int nrows, Store store) {
  constexpr int NS = last_ns(N);
  constexpr int RAD = N / NS;
  constexpr int NB = NS;
  constexpr int TOT = NB * NTR;
  constexpr int ITEMS = (TOT + NT - 1) / NT;
  constexpr int TWS = kTwMax / N;
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const int w = (int)threadIdx.x + it * NT;
    if (ITEMS * NT == TOT || w < TOT) {
      const int col = w % NTR;
      const int j = w / NTR;
      cx<R> v[RAD];
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[r] = s[(j + r * NB) * LD + col];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          cx<R> t = tw[r * j * TWS];
          t.y = -t.y;
          v[r] = cmul<R>(v[r], t);
        }
      }
      dft<R, RAD, true>(v);
#pragma unroll
      for (int r = 0; r < (RAD > 1 ? RAD / 2 : 1); ++r) {
        const int row = j + r * NS;
        if (row < nrows) store(row, col, v[r]);
      }
    }
  }
}

}  // namespace msc
