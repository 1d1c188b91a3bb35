// Sinogram filtering for filtered back projection (filters.py:35-113).
//
// The reference zero-pads each row to L = 2*nfreq, nfreq = 2 * 2^ceil(log2 npix), multiplies
// its FFT by a real, even filter array and keeps the first npix samples of the inverse
// FFT.  For outputs 0..npix-1 that is the linear convolution with the real, even taps
// h[k] = (1/L) sum_j H[j] cos(2 pi j k / L) at lags |k| < npix (no wrap-around: L >= 4 npix).
// The taps are computed once on the host in FP64; the GPU convolves each row with them,
// accumulating in FP64 like numpy's float64 FFT path.
#include <cmath>
#include <vector>

#include "itomo_internal.h"

namespace itomo {

// filters.py:35-76 (dpc = False): half filter for k = 0..nfreq, then mirrored.
static double filter_value(int k, int nfreq, int ftype)
{
    const double f = 2.0 * k / (2.0 * nfreq);   // ramp
    if (k == 0) return 0.0;
    const double w = 2.0 * M_PI * k / (2.0 * nfreq);
    const double d = 1.0;
    switch (ftype) {
    case 1: return f * std::sin(w) / (2.0 * d * 2.0 * d * w);   // shepp-logan
    case 2: return f * std::cos(w) / (2.0 * d * w);             // cosine
    case 3: return f * (0.54 + 0.46 * std::cos(w) / d);        // hamming
    case 4: return f * (1.0 + std::cos(w / d) / 2.0);          // hanning
    default: return f;                                          // ramp
    }
}

__global__ void __launch_bounds__(256)
conv_rows_kernel(const float* __restrict__ in, float* __restrict__ out,
                 const double* __restrict__ taps, int npix)
{
    extern __shared__ unsigned char smem[];
    double* h = (double*)smem;
    float* row = (float*)(h + npix);
    const int64_t r = (int64_t)blockIdx.x;
    in += r * npix;
    out += r * npix;
    for (int k = threadIdx.x; k < npix; k += blockDim.x) {
        h[k] = taps[k];
        row[k] = in[k];
    }
    __syncthreads();
    for (int n = threadIdx.x; n < npix; n += blockDim.x) {
        double acc = 0.0;
        for (int m = 0; m < npix; ++m) {
            const int lag = n >= m ? n - m : m - n;
            acc = fma((double)row[m], h[lag], acc);
        }
        out[n] = (float)acc;
    }
}

}  // namespace itomo

using namespace itomo;

extern "C" {

int itomo_filter_taps(int npix, int ftype, double* taps_host)
{
    if (npix < 1 || npix > 4096 || !taps_host || ftype < 0 || ftype > 5)
        return set_error(ITOMO_EINVAL, "filter_taps: bad argument");
    int p2 = 1;
    while (p2 < npix) p2 <<= 1;
    const int nfreq = 2 * p2, L = 2 * nfreq;
    std::vector<double> H(L);
    for (int j = 0; j < L; ++j) {
        const int k = j <= nfreq ? j : L - j;
        H[j] = ftype == 5 ? 1.0 : filter_value(k, nfreq, ftype);   // 5: 'none'
    }
    for (int k = 0; k < npix; ++k) {
        double s = 0.0;
        for (int j = 0; j < L; ++j)
            s += H[j] * std::cos(2.0 * M_PI * (double)((int64_t)j * k % L) / L);
        taps_host[k] = s / L;
    }
    return ITOMO_OK;
}

int itomo_filter_rows(const float* in, float* out, int nrows, int npix, const double* taps,
                      void* stream)
{
    if (!in || !out || !taps || nrows < 1 || npix < 1 || npix > 4096)
        return set_error(ITOMO_EINVAL, "filter_rows: bad argument");
    const size_t lds = (size_t)npix * (sizeof(double) + sizeof(float));
    conv_rows_kernel<<<nrows, 256, lds, (hipStream_t)stream>>>(in, out, taps, npix);
    return check_hip(hipGetLastError(), "filter_rows launch");
}

}  // extern "C"
