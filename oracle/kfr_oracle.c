/* Test infrastructure (see kfr_oracle.h): CPU restatement of the reference DFT path in long double.
 * Angles are reduced exactly in integers before the long double cos/sin, so the oracle is accurate
 * well beyond FP64 for every size it is used at. */
#include "kfr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef long double ld;
static const ld PI_L = 3.141592653589793238462643383279502884L;

/* exp(-+ 2 pi i a / m) for an exactly reduced integer a */
static void twiddle(uint64_t a, uint64_t m, int inverse, ld* re, ld* im)
{
    a %= m;
    const ld x = 2.0L * PI_L * (ld)a / (ld)m;
    *re        = cosl(x);
    *im        = inverse ? sinl(x) : -sinl(x);
}

/* reference_dft_po2_pass (include/kfr/dft/reference_dft.hpp:43-68): decimation in time, radix 2 */
static void po2_pass(size_t n, int inverse, const ld* in, size_t is, ld* out)
{
    if (n == 1)
    {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    const size_t h = n / 2;
    ld* e          = (ld*)malloc(sizeof(ld) * 2 * h);
    ld* o          = (ld*)malloc(sizeof(ld) * 2 * h);
    po2_pass(h, inverse, in, 2 * is, e);
    po2_pass(h, inverse, in + is, 2 * is, o);
    for (size_t k = 0; k < h; ++k)
    {
        ld wr, wi;
        twiddle(k, n, inverse, &wr, &wi);
        const ld tr = o[2 * k] * wr - o[2 * k + 1] * wi;
        const ld ti = o[2 * k] * wi + o[2 * k + 1] * wr;
        out[2 * k]           = e[2 * k] + tr;
        out[2 * k + 1]       = e[2 * k + 1] + ti;
        out[2 * (k + h)]     = e[2 * k] - tr;
        out[2 * (k + h) + 1] = e[2 * k + 1] - ti;
    }
    free(e);
    free(o);
}

static void dft_ld(size_t n, const ld* in, ld* out, int inverse)
{
    if (n == 0)
        return;
    if ((n & (n - 1)) == 0)
    {
        po2_pass(n, inverse, in, 2, out);
        return;
    }
    /* reference_dft_nonpo2 (reference_dft.hpp:86-113): direct sums */
    ld* wr = (ld*)malloc(sizeof(ld) * n);
    ld* wi = (ld*)malloc(sizeof(ld) * n);
    for (size_t a = 0; a < n; ++a)
        twiddle(a, n, inverse, &wr[a], &wi[a]);
    for (size_t k = 0; k < n; ++k)
    {
        ld sr = 0, si = 0;
        size_t a = 0;
        for (size_t j = 0; j < n; ++j)
        {
            sr += in[2 * j] * wr[a] - in[2 * j + 1] * wi[a];
            si += in[2 * j] * wi[a] + in[2 * j + 1] * wr[a];
            a += k;
            if (a >= n)
                a -= n;
        }
        out[2 * k]     = sr;
        out[2 * k + 1] = si;
    }
    free(wr);
    free(wi);
}

void oracle_dft(size_t n, const double* in, double* out, int inverse)
{
    ld* a = (ld*)malloc(sizeof(ld) * 2 * (n ? n : 1));
    ld* b = (ld*)malloc(sizeof(ld) * 2 * (n ? n : 1));
    for (size_t i = 0; i < 2 * n; ++i)
        a[i] = in[i];
    dft_ld(n, a, b, inverse);
    for (size_t i = 0; i < 2 * n; ++i)
        out[i] = (double)b[i];
    free(a);
    free(b);
}

void oracle_rdft(size_t n, const double* in, double* out, int fmt)
{
    ld* a = (ld*)calloc(2 * n, sizeof(ld));
    ld* b = (ld*)malloc(sizeof(ld) * 2 * n);
    for (size_t i = 0; i < n; ++i)
        a[2 * i] = in[i];
    dft_ld(n, a, b, 0);
    const size_t h = n / 2;
    for (size_t k = 0; k < h; ++k)
    {
        out[2 * k]     = (double)b[2 * k];
        out[2 * k + 1] = (double)b[2 * k + 1];
    }
    if (fmt == 1)
    {
        out[1]         = 0.0;
        out[2 * h]     = (double)b[2 * h];
        out[2 * h + 1] = 0.0;
    }
    else
        out[1] = (double)b[2 * h]; /* Perm: Nyquist real part in X[0].im */
    free(a);
    free(b);
}

void oracle_irdft(size_t n, const double* in, double* out, int fmt)
{
    const size_t h = n / 2;
    ld* a          = (ld*)calloc(2 * n, sizeof(ld));
    ld* b          = (ld*)malloc(sizeof(ld) * 2 * n);
    /* Hermitian extension of the packed half spectrum; imaginary parts of DC and Nyquist dropped */
    a[0]     = in[0];
    a[2 * h] = fmt == 1 ? in[2 * h] : in[1];
    for (size_t k = 1; k < h; ++k)
    {
        a[2 * k]           = in[2 * k];
        a[2 * k + 1]       = in[2 * k + 1];
        a[2 * (n - k)]     = in[2 * k];
        a[2 * (n - k) + 1] = -in[2 * k + 1];
    }
    dft_ld(n, a, b, 1);
    for (size_t i = 0; i < n; ++i)
        out[i] = (double)b[2 * i];
    free(a);
    free(b);
}

void oracle_dct(size_t n, const double* in, double* out, int inverse)
{
    for (size_t k = 0; k < n; ++k)
    {
        ld s = 0;
        for (size_t j = 0; j < n; ++j)
        {
            /* cos(pi k (2j+1) / (2n)) = cos(2 pi a / (4n)), a = k (2j+1) mod 4n */
            ld c, dummy;
            if (!inverse)
            {
                twiddle((uint64_t)k * (2 * j + 1), 4 * n, 0, &c, &dummy);
                s += in[j] * c;
            }
            else
            {
                twiddle((uint64_t)j * (2 * k + 1), 4 * n, 0, &c, &dummy);
                s += (j == 0 ? 0.5L : 1.0L) * in[j] * c;
            }
        }
        out[k] = (double)s;
    }
}

void oracle_convolve(const double* a, size_t na, const double* b, size_t nb, double* out, int correlate, int cplx,
                     int keep_real_only)
{
    const size_t ny = na + nb - 1;
    const size_t w  = cplx ? 2 : 1;
    for (size_t t = 0; t < ny; ++t)
    {
        ld sr = 0, si = 0;
        const size_t j0 = t + 1 > na ? t + 1 - na : 0;
        for (size_t j = j0; j <= t && j < nb; ++j)
        {
            const size_t i  = t - j;
            const size_t jb = correlate ? nb - 1 - j : j;
            if (!cplx)
                sr += (ld)a[i] * b[jb];
            else
            {
                const ld ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * jb], bi = b[2 * jb + 1];
                sr += ar * br - ai * bi;
                si += ar * bi + ai * br;
            }
        }
        out[w * t] = (double)sr;
        if (cplx)
            out[w * t + 1] = keep_real_only ? 0.0 : (double)si;
    }
}

void oracle_chirp(size_t n, double* out)
{
    /* dft-impl.hpp:222-233: k advances by 2m-1 modulo 2n, so k_m = m^2 mod 2n */
    uint64_t k = 0;
    for (size_t m = 0; m < n; ++m)
    {
        if (m > 0)
        {
            k += 2 * m - 1;
            k %= 2 * n;
        }
        ld re, im;
        twiddle(k, 2 * n, 0, &re, &im);
        out[2 * m]     = (double)re;
        out[2 * m + 1] = (double)im;
    }
}

void oracle_bitrev(size_t n, int64_t* perm)
{
    int bits = 0;
    while (((size_t)1 << bits) < n)
        ++bits;
    for (size_t i = 0; i < n; ++i)
    {
        size_t r = 0, v = i;
        for (int b = 0; b < bits; ++b, v >>= 1)
            r = (r << 1) | (v & 1);
        perm[i] = (int64_t)r;
    }
}

void oracle_digitrev(const int* radices, size_t count, int64_t* perm)
{
    /* Output index o has mixed-radix digits d_s (radix r_s, d_0 least significant). The reference's
     * final stage reads transposed, leaving the last digit in place and reversing the others:
     * perm = d_{k-1} n / r_{k-1} + sum_{s<k-1} d_s W_s, W_s = prod_{s<t<k-1} r_t. */
    size_t n = 1;
    for (size_t s = 0; s < count; ++s)
        n *= (size_t)radices[s];
    size_t W[64];
    for (size_t s = 0; s + 1 < count; ++s)
    {
        W[s] = 1;
        for (size_t t = s + 1; t + 1 < count; ++t)
            W[s] *= (size_t)radices[t];
    }
    for (size_t o = 0; o < n; ++o)
    {
        size_t x = o, p = 0;
        for (size_t s = 0; s < count; ++s)
        {
            const size_t d = x % (size_t)radices[s];
            x /= (size_t)radices[s];
            p += s + 1 < count ? d * W[s] : d * (n / (size_t)radices[s]);
        }
        perm[o] = (int64_t)p;
    }
}
