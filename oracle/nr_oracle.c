/*
 * nr_oracle.c -- CPU FP64 oracle for the numrs FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  It is used by tests/ (as the checker), by
 * __graft_entry__.smoke() (as the checker) and by bench.py's cpu_baseline leg
 * (as the timed CPU restatement of the reference path).
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  The
 * reference is Rust, which never contracts a*b+c into an FMA, so contraction is
 * switched off to keep the restatement's rounding sequence equal to the Rust
 * source's.
 *
 * The reference crate cannot be compiled here (no cargo/rustc, FFT modules
 * commented out of lib.rs:14-128, Convolve/Correlation fail to compile), so
 * each routine below restates the Rust source it cites:
 *   - "literal" routines reproduce the reference operation by operation;
 *   - "NR" routines compute what the reference routine is meant to compute
 *     (Numerical Recipes in C, 2nd ed. sections 12.2-12.4, 13.1-13.2) where
 *     the Rust source is defective.  Each divergence is stated at the routine.
 *
 * Parity pins (tests/test_oracle.py): numpy/pocketfft golden vectors in
 * tests/golden/, brute-force DFT, closed forms (delta, constant, pure tone),
 * Parseval, linearity, round trips, separability (fourn), direct O(n*m)
 * circular convolution / correlation, and the reference's own unit-test known
 * answers (FFT_1.rs:245-267, Convolve.rs:346-360/405-424,
 * Correlation.rs:406-502, Fourn.rs:466-476).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Rust's std::f64::consts::PI, the nearest double to pi. */
static const double ORC_PI = 3.14159265358979323846264338327950288;

/* Status codes, equal to the library's NRFFT_* codes in include/nrfft.h. */
#define ORC_OK 0
#define ORC_E_ISIGN (-1)
#define ORC_E_DIMS (-2)
#define ORC_E_NOT_POW2 (-3)
#define ORC_E_LEN (-4)
#define ORC_E_EMPTY (-5)
#define ORC_E_RESP_TOO_LONG (-6)
#define ORC_E_NOMEM (-8)

static int is_pow2(size_t n) { return n != 0 && (n & (n - 1)) == 0; }

/* ------------------------------------------------------------------------ */
/* Seeded inputs.  The reference has no RNG on the FFT path; every seeded    */
/* workload uses this generator (SplitMix64, u = (x >> 11) * 2^-53).         */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64_next(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double u01(uint64_t *s) { return (double)(splitmix64_next(s) >> 11) * 0x1.0p-53; }

/* n draws of uniform[-1,1) = 2u-1 in storage order. */
void orc_gen_uniform(uint64_t seed, size_t n, double *out) {
    uint64_t s = seed;
    for (size_t i = 0; i < n; ++i) out[i] = 2.0 * u01(&s) - 1.0;
}

/* n draws of N(0,1) by Box-Muller, two u per pair:
 * r = sqrt(-2 ln(1-u1)), out[2k] = r cos(2 pi u2), out[2k+1] = r sin(2 pi u2). */
void orc_gen_normal(uint64_t seed, size_t n, double *out) {
    uint64_t s = seed;
    for (size_t i = 0; i < n; i += 2) {
        double u1 = u01(&s), u2 = u01(&s);
        double r = sqrt(-2.0 * log(1.0 - u1));
        double a = 2.0 * ORC_PI * u2;
        out[i] = r * cos(a);
        if (i + 1 < n) out[i + 1] = r * sin(a);
    }
}

/* Generator state after `draws` uniforms from `seed` (lets a workload draw a
 * second block -- e.g. responses after signals -- from the same stream). */
uint64_t orc_gen_skip(uint64_t seed, size_t draws) { return seed + (uint64_t)draws * 0x9E3779B97F4A7C15ull; }

/* ------------------------------------------------------------------------ */
/* four1                                                                     */
/* ------------------------------------------------------------------------ */

/* Literal restatement of FFT_1::four1 (FFT_1.rs:5-44) with its two stage
 * kernels process_butterflies_sequential (FFT_1.rs:46-72, trig recurrence,
 * mmax < 1024) and process_butterflies_parallel (FFT_1.rs:74-107, direct
 * cos/sin table, mmax >= 1024).  The parallel kernel's butterflies are
 * independent, so a serial loop gives the same bits; the table is evaluated
 * with the same operation order: ((isign * 2.0) * PI) * m / mmax.
 * Kernel: exp(+isign * 2 pi i j k / nn); no normalisation.  `data` holds 2*nn
 * doubles (the reference loop bound is data.len(), FFT_1.rs:52-56). */
void orc_four1_literal(double *data, size_t nn, int isign) {
    size_t n = nn * 2;
    size_t j = 1;
    for (size_t i = 1; i < n; i += 2) { /* FFT_1.rs:9-22 */
        if (j > i) {
            double t = data[j - 1]; data[j - 1] = data[i - 1]; data[i - 1] = t;
            t = data[j]; data[j] = data[i]; data[i] = t;
        }
        size_t m = nn;
        while (m >= 2 && j > m) { j -= m; m >>= 1; }
        j += m;
    }
    size_t mmax = 2;
    while (n > mmax) { /* FFT_1.rs:26-43 */
        size_t istep = mmax << 1;
        double theta = (double)isign * (2.0 * ORC_PI / (double)mmax);
        double wtemp = sin(0.5 * theta);
        double wpr = -2.0 * wtemp * wtemp;
        double wpi = sin(theta);
        if (mmax >= 1024) { /* FFT_1.rs:74-107 */
            size_t half = mmax / 2;
            double *wr_t = (double *)malloc(half * sizeof(double));
            double *wi_t = (double *)malloc(half * sizeof(double));
            for (size_t m = 0; m < half; ++m) {
                double angle = (double)isign * 2.0 * ORC_PI * (double)m / (double)mmax;
                wr_t[m] = cos(angle);
                wi_t[m] = sin(angle);
            }
            for (size_t c = 0; c + istep <= n; c += istep) {
                double *chunk = data + c;
                for (size_t m = 0; m < half; ++m) {
                    double wr = wr_t[m], wi = wi_t[m];
                    for (size_t i = m * 2; i < istep - mmax; i += istep) {
                        size_t jj = i + mmax;
                        double tempr = wr * chunk[jj] - wi * chunk[jj + 1];
                        double tempi = wr * chunk[jj + 1] + wi * chunk[jj];
                        chunk[jj] = chunk[i] - tempr;
                        chunk[jj + 1] = chunk[i + 1] - tempi;
                        chunk[i] += tempr;
                        chunk[i + 1] += tempi;
                    }
                }
            }
            free(wr_t);
            free(wi_t);
        } else { /* FFT_1.rs:46-72 */
            double wr = 1.0, wi = 0.0;
            for (size_t m = 0; m < mmax; m += 2) {
                for (size_t i = m; i < n; i += istep) {
                    size_t jj = i + mmax;
                    if (jj >= n) continue;
                    double tempr = wr * data[jj] - wi * data[jj + 1];
                    double tempi = wr * data[jj + 1] + wi * data[jj];
                    data[jj] = data[i] - tempr;
                    data[jj + 1] = data[i + 1] - tempi;
                    data[i] += tempr;
                    data[i + 1] += tempi;
                }
                double wt = wr;
                wr = wt * wpr - wi * wpi + wr;
                wi = wi * wpr + wt * wpi + wi;
            }
        }
        mmax = istep;
    }
}

/* Restatement of Real_FT::four1 (Real_FT.rs:429-476): NR four1 with the trig
 * recurrence in every stage.  This is the FFT inside realft/convlv/correl. */
void orc_four1_nr(double *data, size_t nn, int isign) {
    size_t n = nn * 2;
    size_t j = 1;
    for (size_t i = 1; i < n; i += 2) {
        if (j > i) {
            double t = data[j - 1]; data[j - 1] = data[i - 1]; data[i - 1] = t;
            t = data[j]; data[j] = data[i]; data[i] = t;
        }
        size_t m = nn;
        while (m >= 2 && j > m) { j -= m; m >>= 1; }
        j += m;
    }
    size_t mmax = 2;
    while (n > mmax) {
        size_t istep = mmax << 1;
        double theta = (double)isign * 2.0 * ORC_PI / (double)mmax;
        double wtemp = sin(theta / 2.0);
        double wpr = -2.0 * wtemp * wtemp;
        double wpi = sin(theta);
        double wr = 1.0, wi = 0.0;
        for (size_t m = 1; m < mmax; m += 2) {
            for (size_t i = m; i <= n; i += istep) {
                size_t jj = i + mmax;
                double tempr = wr * data[jj - 1] - wi * data[jj];
                double tempi = wr * data[jj] + wi * data[jj - 1];
                data[jj - 1] = data[i - 1] - tempr;
                data[jj] = data[i] - tempi;
                data[i - 1] += tempr;
                data[i] += tempi;
            }
            double wt = wr;
            wr = wt * wpr - wi * wpi + wr;
            wi = wi * wpr + wt * wpi + wi;
        }
        mmax = istep;
    }
}

/* Brute-force DFT in long double (pin for small n): out[k] = sum_j x[j] e^{isign 2 pi i jk/n}. */
void orc_dft_brute(const double *in, size_t nn, int isign, double *out) {
    const long double pi = 3.141592653589793238462643383279502884L;
    for (size_t k = 0; k < nn; ++k) {
        long double sr = 0.0L, si = 0.0L;
        for (size_t j = 0; j < nn; ++j) {
            size_t m = (j * k) % nn;
            long double a = (long double)isign * 2.0L * pi * (long double)m / (long double)nn;
            long double c = cosl(a), s = sinl(a);
            sr += (long double)in[2 * j] * c - (long double)in[2 * j + 1] * s;
            si += (long double)in[2 * j] * s + (long double)in[2 * j + 1] * c;
        }
        out[2 * k] = (double)sr;
        out[2 * k + 1] = (double)si;
    }
}

/* ------------------------------------------------------------------------ */
/* realft                                                                    */
/* ------------------------------------------------------------------------ */

/* NR-semantic realft (Real_FT.rs:4-21 with its untangle loops 23-213 read with
 * NR's 1-based indices).  Forward (isign=+1): n real samples -> packed
 * [F_0, F_{n/2}, Re F_1, Im F_1, ..., Re F_{n/2-1}, Im F_{n/2-1}] with
 * F_k = sum_j x_j e^{+2 pi i jk/n}.  Inverse (isign=-1): packed -> (n/2) x.
 *
 * Divergence from the Rust source: Real_FT.rs:61-64 (and 102-105, 157-160,
 * 198-201) keep NR's 1-based index formulas i1=2i-1 .. i4=i3+1 but index a
 * 0-based slice, and the n>=1024 branch uses twiddle w^(i-2) instead of
 * w^(i-1) (Real_FT.rs:93,189).  Here i1..i4 are shifted to 0-based and the
 * twiddle recurrence is NR's, so the output is the real DFT the routine is
 * documented to produce.  orc_realft_literal below keeps the defect. */
void orc_realft(double *data, size_t n, int isign) {
    double c1 = 0.5, c2, theta;
    theta = ORC_PI / (double)(n >> 1);
    if (isign == 1) {
        c2 = -0.5;
        orc_four1_nr(data, n >> 1, 1);
    } else {
        c2 = 0.5;
        theta = -theta;
    }
    double wtemp = sin(0.5 * theta);
    double wpr = -2.0 * wtemp * wtemp;
    double wpi = sin(theta);
    double wr = 1.0 + wpr, wi = wpi;
    size_t np3 = n + 3;
    for (size_t i = 2; i <= (n >> 2); ++i) {
        /* NR's 1-based positions, shifted to 0-based by the D() accessor */
        size_t i1 = i + i - 1, i2 = 1 + i1, i3 = np3 - i2, i4 = 1 + i3;
#define D(k) data[(k) - 1]
        double h1r = c1 * (D(i1) + D(i3));
        double h1i = c1 * (D(i2) - D(i4));
        double h2r = -c2 * (D(i2) + D(i4));
        double h2i = c2 * (D(i1) - D(i3));
        D(i1) = h1r + wr * h2r - wi * h2i;
        D(i2) = h1i + wr * h2i + wi * h2r;
        D(i3) = h1r - wr * h2r + wi * h2i;
        D(i4) = -h1i + wr * h2i + wi * h2r;
#undef D
        double wt = wr;
        wr = wt * wpr - wi * wpi + wr;
        wi = wi * wpr + wt * wpi + wi;
    }
    if (isign == 1) {
        double h1r = data[0];
        data[0] = h1r + data[1];
        data[1] = h1r - data[1];
    } else {
        double h1r = data[0];
        data[0] = c1 * (h1r + data[1]);
        data[1] = c1 * (h1r - data[1]);
        orc_four1_nr(data, n >> 1, -1);
    }
}

/* Literal restatement of Real_FT::realft (Real_FT.rs:4-213), defects kept:
 * 1-based index formulas on a 0-based slice, and for n >= 1024 the direct
 * twiddle table with exponent (i-2).  Needs data to hold at least n+1 values
 * (it writes data[n]; the reference callers pass 2n, Convolve.rs:81).
 * Documentation only: this is not a parity target. */
void orc_realft_literal(double *data, size_t n, int isign) {
    size_t np3 = n + 3, q = n / 4;
    double c1 = 0.5, c2 = (isign == 1) ? -0.5 : 0.5;
    double theta = (isign == 1 ? ORC_PI : -ORC_PI) / (double)(n / 2);
    if (isign == 1) orc_four1_nr(data, n / 2, 1);
    if (isign != 1) {
        double h1r = data[0];
        data[0] = c1 * (h1r + data[1]);
        data[1] = c1 * (h1r - data[1]);
    }
    double wtemp = sin(0.5 * theta);
    double wpr = -2.0 * wtemp * wtemp;
    double wpi = sin(theta);
    double wr = 1.0 + wpr, wi = wpi;
    for (size_t i = 2; i <= q; ++i) {
        size_t i1 = 2 * i - 1, i2 = i1 + 1, i3 = np3 - i2, i4 = i3 + 1;
        double twr = wr, twi = wi;
        if (n >= 1024) {
            double angle = (double)(i - 2) * (theta);
            twr = cos(angle);
            twi = sin(angle);
        }
        double h1r = c1 * (data[i1] + data[i3]);
        double h1i = c1 * (data[i2] - data[i4]);
        double h2r = -c2 * (data[i2] + data[i4]);
        double h2i = c2 * (data[i1] - data[i3]);
        data[i1] = h1r + twr * h2r - twi * h2i;
        data[i2] = h1i + twr * h2i + twi * h2r;
        data[i3] = h1r - twr * h2r + twi * h2i;
        data[i4] = -h1i + twr * h2i + twi * h2r;
        double wt = wr;
        wr = wt * wpr - wi * wpi + wr;
        wi = wi * wpr + wt * wpi + wi;
    }
    if (isign == 1) {
        double h1r = data[0];
        data[0] = h1r + data[1];
        data[1] = h1r - data[1];
    } else {
        orc_four1_nr(data, n / 2, -1);
    }
}

/* ------------------------------------------------------------------------ */
/* fourn                                                                     */
/* ------------------------------------------------------------------------ */

/* NR fourn: in-place unnormalised n-D DFT, C order (nn[0] slowest, last index
 * fastest), kernel e^{+isign 2 pi i}.  The reference has no in-memory fourn
 * (Fourn.rs:56 takes no data); the calling convention is the one of its only
 * caller, Real_FT3.rs:33-35: Fourn(&mut flat, &[nn1, nn2, nn3/2], 3, isign).
 * Validation follows Fourn.rs:367-378 (ndim==0 or ndim>len -> dims error;
 * isign not +-1 -> isign error; any nn<=1 -> dims error) and this build
 * additionally rejects non-powers of two.
 * Algorithm: per dimension (fastest first), a bit-reversal over that index
 * followed by radix-2 Danielson-Lanczos stages with the trig recurrence,
 * strided by the product of the faster dimensions. */
int orc_fourn(double *data, const size_t *nn, size_t ndim, int isign) {
    if (ndim == 0) return ORC_E_DIMS;
    if (isign != 1 && isign != -1) return ORC_E_ISIGN;
    for (size_t d = 0; d < ndim; ++d) {
        if (nn[d] <= 1) return ORC_E_DIMS;
        if (!is_pow2(nn[d])) return ORC_E_NOT_POW2;
    }
    size_t ntot = 1;
    for (size_t d = 0; d < ndim; ++d) ntot *= nn[d];
#define dd(k) data[(k) - 1] /* NR's 1-based positions */
    size_t nprev = 1;
    for (size_t idim = ndim; idim >= 1; --idim) {
        size_t n = nn[idim - 1];
        size_t nrem = ntot / (n * nprev);
        size_t ip1 = nprev << 1;
        size_t ip2 = ip1 * n;
        size_t ip3 = ip2 * nrem;
        size_t i2rev = 1;
        for (size_t i2 = 1; i2 <= ip2; i2 += ip1) {
            if (i2 < i2rev) {
                for (size_t i1 = i2; i1 <= i2 + ip1 - 2; i1 += 2) {
                    for (size_t i3 = i1; i3 <= ip3; i3 += ip2) {
                        size_t i3rev = i2rev + i3 - i2;
                        double t = dd(i3); dd(i3) = dd(i3rev); dd(i3rev) = t;
                        t = dd(i3 + 1); dd(i3 + 1) = dd(i3rev + 1); dd(i3rev + 1) = t;
                    }
                }
            }
            size_t ibit = ip2 >> 1;
            while (ibit >= ip1 && i2rev > ibit) { i2rev -= ibit; ibit >>= 1; }
            i2rev += ibit;
        }
        size_t ifp1 = ip1;
        while (ifp1 < ip2) {
            size_t ifp2 = ifp1 << 1;
            double theta = (double)isign * 2.0 * ORC_PI / (double)(ifp2 / ip1);
            double wtemp = sin(0.5 * theta);
            double wpr = -2.0 * wtemp * wtemp;
            double wpi = sin(theta);
            double wr = 1.0, wi = 0.0;
            for (size_t i3 = 1; i3 <= ifp1; i3 += ip1) {
                for (size_t i1 = i3; i1 <= i3 + ip1 - 2; i1 += 2) {
                    for (size_t i2 = i1; i2 <= ip3; i2 += ifp2) {
                        size_t k1 = i2, k2 = k1 + ifp1;
                        double tempr = wr * dd(k2) - wi * dd(k2 + 1);
                        double tempi = wr * dd(k2 + 1) + wi * dd(k2);
                        dd(k2) = dd(k1) - tempr;
                        dd(k2 + 1) = dd(k1 + 1) - tempi;
                        dd(k1) += tempr;
                        dd(k1 + 1) += tempi;
                    }
                }
                double wt = wr;
                wr = wt * wpr - wi * wpi + wr;
                wi = wi * wpr + wt * wpi + wi;
            }
            ifp1 = ifp2;
        }
        nprev *= n;
        if (idim == 1) break;
    }
#undef dd
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* convlv                                                                    */
/* ------------------------------------------------------------------------ */

/* Response placement of Convolve::prepare_response_function
 * (Convolve.rs:41-63), reproduced exactly, mid = (m-1)/2:
 *   [0, mid)      <- respns[m-mid+i]
 *   [mid, m)      <- respns[i]
 *   [m, n-mid)    <- 0
 *   [n-mid, n)    <- respns[i-(n-mid)]                                      */
void orc_prepare_response(double *padded, const double *respns, size_t n, size_t m) {
    size_t mid = (m - 1) / 2;
    for (size_t i = 0; i < n; ++i) {
        if (i < mid) padded[i] = respns[m - mid + i];
        else if (i < m) padded[i] = respns[i];
        else if (i < n - mid) padded[i] = 0.0;
        else padded[i] = respns[i - (n - mid)];
    }
}

/* NR-semantic convlv (Convolve.rs:8-38): validate (Convolve.rs:13-21: empty ->
 * EmptyInput, m>n -> ResponseTooLong, isign not +-1 -> InvalidIsign; this
 * build also requires n to be a power of two), place the response, forward
 * realft both, multiply (isign=+1) or divide with the |R|^2 < 1e-12 -> 0 guard
 * (isign=-1, Convolve.rs:118-122) on every bin including DC and Nyquist,
 * scale by 1/(n/2) (Convolve.rs:114-115), inverse realft, return n values.
 * Divergence: the Rust spectral loop (Convolve.rs:106,134-135) skips DC and
 * Nyquist of the packed spectrum; here they are processed as in NR convlv, so
 * the result is the exact circular convolution. */
int orc_convlv(const double *data, size_t n, const double *respns, size_t m, int isign, double *ans) {
    if (n == 0 || m == 0) return ORC_E_EMPTY;
    if (m > n) return ORC_E_RESP_TOO_LONG;
    if (isign != 1 && isign != -1) return ORC_E_ISIGN;
    if (!is_pow2(n) || n < 2) return ORC_E_NOT_POW2;
    double *fd = (double *)calloc(n, sizeof(double));
    double *fr = (double *)calloc(n, sizeof(double));
    if (!fd || !fr) { free(fd); free(fr); return ORC_E_NOMEM; }
    memcpy(fd, data, n * sizeof(double));
    orc_prepare_response(fr, respns, n, m);
    orc_realft(fd, n, 1);
    orc_realft(fr, n, 1);
    double no2 = (double)(n >> 1);
    for (size_t k = 0; k < 2; ++k) { /* DC (index 0) and Nyquist (index 1) are real */
        if (isign == 1) ans[k] = fd[k] * fr[k] / no2;
        else {
            double mag2 = fr[k] * fr[k];
            ans[k] = (mag2 < 1e-12) ? 0.0 : fd[k] * fr[k] / mag2 / no2;
        }
    }
    for (size_t i = 2; i < n; i += 2) {
        double dr = fd[i], di = fd[i + 1], rr = fr[i], ri = fr[i + 1];
        if (isign == 1) {
            ans[i] = (dr * rr - di * ri) / no2;
            ans[i + 1] = (dr * ri + di * rr) / no2;
        } else {
            double mag2 = rr * rr + ri * ri;
            if (mag2 < 1e-12) { ans[i] = 0.0; ans[i + 1] = 0.0; }
            else {
                ans[i] = (dr * rr + di * ri) / mag2 / no2;
                ans[i + 1] = (di * rr - dr * ri) / mag2 / no2;
            }
        }
    }
    orc_realft(ans, n, -1);
    free(fd);
    free(fr);
    return ORC_OK;
}

/* Direct circular convolution with the same placement (pin for orc_convlv):
 * ans[k] = sum_j data[(k - j) mod n] * padded[j], over the non-zeros of padded. */
int orc_convlv_direct(const double *data, size_t n, const double *respns, size_t m, double *ans) {
    if (n == 0 || m == 0) return ORC_E_EMPTY;
    if (m > n) return ORC_E_RESP_TOO_LONG;
    double *p = (double *)malloc(n * sizeof(double));
    if (!p) return ORC_E_NOMEM;
    orc_prepare_response(p, respns, n, m);
    for (size_t k = 0; k < n; ++k) {
        long double s = 0.0L;
        for (size_t j = 0; j < n; ++j)
            if (p[j] != 0.0) s += (long double)data[(k + n - j) % n] * (long double)p[j];
        ans[k] = (double)s;
    }
    free(p);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* correl                                                                    */
/* ------------------------------------------------------------------------ */

/* Correlation::correl (Correlation.rs:8-34).  n <= 32: the reference's direct
 * linear formula (Correlation.rs:37-50) r[l] = sum_{i<n-l} d1[i+l] d2[i].
 * n > 32: NR correl, the circular correlation r[l] = sum_j d1[(j+l) mod n] d2[j]
 * (negative lags wrapped to the top), by realft of both, F1 conj(F2)/(n/2),
 * inverse realft.  Divergence: the Rust n>32 branch calls a placeholder O(n^2)
 * DFT (Correlation.rs:325-386) and reads every second output
 * (Correlation.rs:181-183); n must be a power of two here. */
int orc_correl(const double *d1, const double *d2, size_t n, double *ans) {
    if (n == 0) return ORC_E_EMPTY;
    if (n <= 32) {
        for (size_t lag = 0; lag < n; ++lag) {
            double s = 0.0;
            for (size_t i = 0; i < n - lag; ++i) s += d1[i + lag] * d2[i];
            ans[lag] = s;
        }
        return ORC_OK;
    }
    if (!is_pow2(n)) return ORC_E_NOT_POW2;
    double *f1 = (double *)malloc(n * sizeof(double));
    double *f2 = (double *)malloc(n * sizeof(double));
    if (!f1 || !f2) { free(f1); free(f2); return ORC_E_NOMEM; }
    memcpy(f1, d1, n * sizeof(double));
    memcpy(f2, d2, n * sizeof(double));
    orc_realft(f1, n, 1);
    orc_realft(f2, n, 1);
    double no2 = (double)(n >> 1);
    ans[0] = f1[0] * f2[0] / no2;
    ans[1] = f1[1] * f2[1] / no2;
    for (size_t i = 2; i < n; i += 2) {
        double ar = f1[i], ai = f1[i + 1], br = f2[i], bi = f2[i + 1];
        ans[i] = (ar * br + ai * bi) / no2;
        ans[i + 1] = (ai * br - ar * bi) / no2;
    }
    orc_realft(ans, n, -1);
    free(f1);
    free(f2);
    return ORC_OK;
}

/* Direct circular correlation (pin for the n > 32 branch). */
void orc_correl_direct_circular(const double *d1, const double *d2, size_t n, double *ans) {
    for (size_t l = 0; l < n; ++l) {
        long double s = 0.0L;
        for (size_t j = 0; j < n; ++j) s += (long double)d1[(j + l) % n] * (long double)d2[j];
        ans[l] = (double)s;
    }
}

/* ------------------------------------------------------------------------ */
/* CPU baseline drivers (bench.py's cpu_baseline leg).                       */
/* Batches run one transform per OpenMP thread, as FFTProcessor::fft_batch   */
/* (FFT_1.rs:185-189) and convlv_batch (Convolve.rs:241-250) do with rayon.  */
/* ------------------------------------------------------------------------ */

/* Returns the number of OpenMP threads used. */
/* ------------------------------------------------------------------------
 * twofft: FFT_2.rs:3-17 (pack), 52-132 (separation) -- NR twofft.
 * NR: the two real inputs packed as z_j = d1_j + i d2_j, one n-point four1
 * (isign +1), then F1[k] = (Z_k + conj Z_{n-k})/2, F2[k] = (Z_k - conj Z_{n-k})/(2i).
 * Divergences of the Rust source corrected here: it runs four1 over a
 * 2n+2-long buffer (four1's loop bound is data.len(), FFT_1.rs:52-56) and
 * pairs bin k with index 2n+2-2k (NR's 1-based mirror, FFT_2.rs:76-78), which
 * is bin n+1-k in 0-based storage.  fft1 / fft2 receive bins 0..n-1 (2n
 * values each).
 * ------------------------------------------------------------------------ */
void orc_twofft(const double *d1, const double *d2, size_t n, double *fft1, double *fft2) {
    for (size_t j = 0; j < n; ++j) {
        fft1[2 * j] = d1[j];
        fft1[2 * j + 1] = d2[j];
    }
    orc_four1_nr(fft1, n, 1);
    fft2[0] = fft1[1];
    fft1[1] = fft2[1] = 0.0;
    for (size_t k = 1; k <= n / 2; ++k) {
        const size_t r = n - k;
        const double rep = 0.5 * (fft1[2 * k] + fft1[2 * r]);
        const double rem = 0.5 * (fft1[2 * k] - fft1[2 * r]);
        const double aip = 0.5 * (fft1[2 * k + 1] + fft1[2 * r + 1]);
        const double aim = 0.5 * (fft1[2 * k + 1] - fft1[2 * r + 1]);
        fft1[2 * k] = rep;
        fft1[2 * k + 1] = aim;
        fft1[2 * r] = rep;
        fft1[2 * r + 1] = -aim;
        fft2[2 * k] = aip;
        fft2[2 * k + 1] = -rem;
        fft2[2 * r] = aip;
        fft2[2 * r + 1] = rem;
    }
}

/* ------------------------------------------------------------------------
 * cosft1: Cos_FT.rs:7-68 -- NR cosft1 (DCT-I), 0-based: y[0..n] (n + 1
 * values) becomes Y_k = (y_0 + (-1)^k y_n)/2 + sum_{j=1}^{n-1} y_j cos(pi jk/n).
 * Pre-pass with recurrence twiddles, realft(n, +1), running sum over the odd
 * slots.  Divergences of the Rust source corrected here: 1-based indices in a
 * 0-based slice (Cos_FT.rs:16-17), realft left unimplemented (Cos_FT.rs:71-75),
 * and the missing y[2] = sum step of NR after realft (Cos_FT.rs:58-64).
 * Accuracy: the twiddles come from NR's trigonometric recurrence, which drifts
 * by ~n eps: 3e-14 of the output at n = 2^14, 7e-13 at 2^17 and 2e-12 at 2^20
 * (relative L2 against scipy's DCT).  cosft2 below drifts the same way
 * (1e-12 at 2^17).
 * ------------------------------------------------------------------------ */
void orc_cosft1(double *y, size_t n) {
    const double theta = ORC_PI / (double)n;
    double wtemp = sin(0.5 * theta);
    const double wpr = -2.0 * wtemp * wtemp, wpi = sin(theta);
    double wr = 1.0, wi = 0.0;
    double sum = 0.5 * (y[0] - y[n]);
    y[0] = 0.5 * (y[0] + y[n]);
    for (size_t j = 1; j < n / 2; ++j) {
        wtemp = wr;
        wr = wtemp * wpr - wi * wpi + wr;
        wi = wi * wpr + wtemp * wpi + wi;
        const double y1 = 0.5 * (y[j] + y[n - j]);
        const double y2 = y[j] - y[n - j];
        y[j] = y1 - wi * y2;
        y[n - j] = y1 + wi * y2;
        sum += wr * y2;
    }
    orc_realft(y, n, 1);
    y[n] = y[1];
    y[1] = sum;
    for (size_t j = 3; j < n; j += 2) {
        sum += y[j];
        y[j] = sum;
    }
}

/* ------------------------------------------------------------------------
 * cosft2: Cos_FT2.rs:7-173 -- NR cosft2, 0-based y[0..n-1].
 *   isign = +1: Y_k = sum_j y_j cos(pi k (j + 1/2) / n)            (DCT-II)
 *   isign = -1: y_j = Y_0/2 + sum_{k>=1} Y_k cos(pi k (j + 1/2) / n)  (DCT-III,
 *               so inverse(forward(y)) = (n/2) y)
 * Divergences of the Rust source corrected here: 1-based indices in a
 * 0-based slice, the pre-pass twiddle wi1 never advanced (Cos_FT2.rs:25-35),
 * recurrences adding wr1/wi1 where NR adds wr/wi (Cos_FT2.rs:56-57, 106-107),
 * the inverse rotation's operands swapped (Cos_FT2.rs:118-119), realft left
 * unimplemented.
 * ------------------------------------------------------------------------ */
void orc_cosft2(double *y, size_t n, int isign) {
    const double theta = 0.5 * ORC_PI / (double)n;
    double wr = 1.0, wi = 0.0, wr1 = cos(theta), wi1 = sin(theta);
    const double wpr = -2.0 * wi1 * wi1, wpi = sin(2.0 * theta);
    double wtemp;
    if (isign == 1) {
        for (size_t i = 0; i < n / 2; ++i) {
            const double y1 = 0.5 * (y[i] + y[n - 1 - i]);
            const double y2 = wi1 * (y[i] - y[n - 1 - i]);
            y[i] = y1 + y2;
            y[n - 1 - i] = y1 - y2;
            wtemp = wr1;
            wr1 = wtemp * wpr - wi1 * wpi + wr1;
            wi1 = wi1 * wpr + wtemp * wpi + wi1;
        }
        orc_realft(y, n, 1);
        for (size_t i = 2; i < n; i += 2) {
            wtemp = wr;
            wr = wtemp * wpr - wi * wpi + wr;
            wi = wi * wpr + wtemp * wpi + wi;
            const double y1 = y[i] * wr - y[i + 1] * wi;
            const double y2 = y[i + 1] * wr + y[i] * wi;
            y[i] = y1;
            y[i + 1] = y2;
        }
        double sum = 0.5 * y[1];
        for (size_t i = n; i >= 2; i -= 2) {
            const double sum1 = sum;
            sum += y[i - 1];
            y[i - 1] = sum1;
        }
    } else {
        const double ytemp = y[n - 1];
        for (size_t i = n; i >= 4; i -= 2) y[i - 1] = y[i - 3] - y[i - 1];
        y[1] = 2.0 * ytemp;
        for (size_t i = 2; i < n; i += 2) {
            wtemp = wr;
            wr = wtemp * wpr - wi * wpi + wr;
            wi = wi * wpr + wtemp * wpi + wi;
            const double y1 = y[i] * wr + y[i + 1] * wi;
            const double y2 = y[i + 1] * wr - y[i] * wi;
            y[i] = y1;
            y[i + 1] = y2;
        }
        orc_realft(y, n, -1);
        for (size_t i = 0; i < n / 2; ++i) {
            const double y1 = y[i] + y[n - 1 - i];
            const double y2 = (0.5 / wi1) * (y[i] - y[n - 1 - i]);
            y[i] = 0.5 * (y1 + y2);
            y[n - 1 - i] = 0.5 * (y1 - y2);
            wtemp = wr1;
            wr1 = wtemp * wpr - wi1 * wpi + wr1;
            wi1 = wi1 * wpr + wtemp * wpi + wi1;
        }
    }
}

/* ------------------------------------------------------------------------
 * rlft3: Real_FT3.rs:8-141 -- NR rlft3, C order, 0-based.
 * data: nn1 x nn2 x nn3 reals; speq: nn1 x 2*nn2 reals (the nn3/2 plane).
 * isign = +1: data <- F[k1][k2][k3] for k3 < nn3/2 (complex, interleaved),
 *             speq <- F[k1][k2][nn3/2],  F = sum x e^{+2 pi i (j.k / nn)}.
 * isign = -1: the inverse; data <- (nn1 nn2 nn3 / 2) x.
 * fourn over (nn1, nn2, nn3/2) runs with length-1 axes dropped (NR's 2-D use
 * sets nn1 = 1).  Divergence of the Rust source corrected here: its forward
 * fill of speq uses the mirrored column (Real_FT3.rs:44-49) where NR fills it
 * in order, so the main loop's mirrored read (Real_FT3.rs:72-77) pairs the
 * wrong bins.  Returns fourn's status (0 = ok).
 * ------------------------------------------------------------------------ */
int orc_rlft3(double *data, double *speq, size_t nn1, size_t nn2, size_t nn3, int isign) {
    const double c1 = 0.5, c2 = -0.5 * isign;
    const double theta = isign * (2.0 * ORC_PI / (double)nn3);
    double wtemp = sin(0.5 * theta);
    const double wpr = -2.0 * wtemp * wtemp, wpi = sin(theta);
    const size_t h = nn3 >> 1;
    size_t dims[3], nd = 0;
    if (nn1 > 1) dims[nd++] = nn1;
    if (nn2 > 1) dims[nd++] = nn2;
    if (h > 1) dims[nd++] = h;
#define RD(a, b, c) data[((a) * nn2 + (b)) * nn3 + (c)]
#define RS(a, c) speq[(a) * 2 * nn2 + (c)]
    if (isign == 1) {
        if (nd) {
            int rc = orc_fourn(data, dims, nd, 1);
            if (rc) return rc;
        }
        for (size_t i1 = 0; i1 < nn1; ++i1)
            for (size_t i2 = 0; i2 < nn2; ++i2) {
                RS(i1, 2 * i2) = RD(i1, i2, 0);
                RS(i1, 2 * i2 + 1) = RD(i1, i2, 1);
            }
    }
    for (size_t i1 = 0; i1 < nn1; ++i1) {
        const size_t j1 = (nn1 - i1) % nn1;
        double wr = 1.0, wi = 0.0;
        for (size_t k3 = 0; k3 <= (nn3 >> 2); ++k3) {
            for (size_t i2 = 0; i2 < nn2; ++i2) {
                const size_t j2 = (nn2 - i2) % nn2;
                if (k3 == 0) {
                    const double h1r = c1 * (RD(i1, i2, 0) + RS(j1, 2 * j2));
                    const double h1i = c1 * (RD(i1, i2, 1) - RS(j1, 2 * j2 + 1));
                    const double h2i = c2 * (RD(i1, i2, 0) - RS(j1, 2 * j2));
                    const double h2r = -c2 * (RD(i1, i2, 1) + RS(j1, 2 * j2 + 1));
                    RD(i1, i2, 0) = h1r + h2r;
                    RD(i1, i2, 1) = h1i + h2i;
                    RS(j1, 2 * j2) = h1r - h2r;
                    RS(j1, 2 * j2 + 1) = h2i - h1i;
                } else {
                    const size_t a = 2 * k3, b = 2 * (h - k3);
                    const double h1r = c1 * (RD(i1, i2, a) + RD(j1, j2, b));
                    const double h1i = c1 * (RD(i1, i2, a + 1) - RD(j1, j2, b + 1));
                    const double h2i = c2 * (RD(i1, i2, a) - RD(j1, j2, b));
                    const double h2r = -c2 * (RD(i1, i2, a + 1) + RD(j1, j2, b + 1));
                    RD(i1, i2, a) = h1r + wr * h2r - wi * h2i;
                    RD(i1, i2, a + 1) = h1i + wr * h2i + wi * h2r;
                    RD(j1, j2, b) = h1r - wr * h2r + wi * h2i;
                    RD(j1, j2, b + 1) = -h1i + wr * h2i + wi * h2r;
                }
            }
            wtemp = wr;
            wr = wtemp * wpr - wi * wpi + wr;
            wi = wi * wpr + wtemp * wpi + wi;
        }
    }
#undef RD
#undef RS
    if (isign == -1 && nd) return orc_fourn(data, dims, nd, -1);
    return 0;
}

int orc_four1_batch(double *data, size_t nn, size_t batch, int isign, int nthreads) {
    int used = 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 1)
        for (long b = 0; b < (long)batch; ++b) orc_four1_literal(data + (size_t)b * 2 * nn, nn, isign);
    }
#else
    (void)nthreads;
    for (size_t b = 0; b < batch; ++b) orc_four1_literal(data + b * 2 * nn, nn, isign);
#endif
    return used;
}

int orc_convlv_batch(const double *data, size_t n, const double *respns, size_t m, size_t batch, int isign,
                     double *ans, int nthreads) {
    int used = 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 1)
        for (long b = 0; b < (long)batch; ++b)
            orc_convlv(data + (size_t)b * n, n, respns + (size_t)b * m, m, isign, ans + (size_t)b * n);
    }
#else
    (void)nthreads;
    for (size_t b = 0; b < batch; ++b) orc_convlv(data + b * n, n, respns + b * m, m, isign, ans + b * n);
#endif
    return used;
}

/* CPU baseline of an uneven batch (FFTProcessor::fft_batch, FFT_1.rs:185-189,
 * parallel over items): item i (ndim[i] = 1 or 2, dims dims[2i], dims[2i+1])
 * at complex offset offs[i] of data, 1-D items through the literal four1,
 * 2-D items through fourn.  Returns the threads used. */
int orc_uneven_batch(double *data, const int *ndim, const size_t *dims, const size_t *offs, size_t count, int isign,
                     int nthreads) {
    int used = 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 1)
        for (long i = 0; i < (long)count; ++i) {
            double *d = data + 2 * offs[i];
            if (ndim[i] == 1) orc_four1_literal(d, dims[2 * i], isign);
            else (void)orc_fourn(d, dims + 2 * i, 2, isign);
        }
    }
#else
    (void)nthreads;
    for (size_t i = 0; i < count; ++i) {
        double *d = data + 2 * offs[i];
        if (ndim[i] == 1) orc_four1_literal(d, dims[2 * i], isign);
        else (void)orc_fourn(d, dims + 2 * i, 2, isign);
    }
#endif
    return used;
}

int orc_fourn_threads(double *data, const size_t *nn, size_t ndim, int isign, int nthreads) {
    /* fourn itself is serial (NR); threads only matter for the batch drivers. */
    (void)nthreads;
    return orc_fourn(data, nn, ndim, isign);
}
