/*
 * ske_oracle.c -- CPU restatement of the reference skewEngine CPU path
 * (reference: src/skewEngine.cuh, src/kernel{SDEM,RADN,BLUR,UNIT}.h,
 * src/helper.h, src/main.cu, src/main.h).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
 * The product library never links or loads this file.
 *
 * Build: -O3 -fopenmp -ffp-contract=off, no -march, no fast-math, i.e. the
 * same floating-point semantics as the reference's generic x86-64 build
 * (CMakeLists.txt:8-9,104): every float/double operation is rounded
 * separately, no fused multiply-add.
 */
#include "ske_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* skewEngine.cuh:465 -- fAngle = atan2(dimy,dimx)*180/M_PI stored as float. */
float orc_fangle(int dimx, int dimy)
{
    return (float)(atan2((double)dimy, (double)dimx) * 180 / M_PI);
}

/* skewEngine.cuh:503-539 (computeParams). */
void orc_compute_params(int dimx, int dimy, int angle, orc_params *p, int *target, float *weight)
{
    const double torads = M_PI / 180.0; /* skewEngine.cuh:258 */
    float fAngle = orc_fangle(dimx, dimy);
    int fA = (int)fAngle;
    p->fAngle = fAngle;
    p->dim_l = dimx > dimy ? dimx : dimy; /* :475 */
    p->a = angle;
    p->isT = angle > fA && angle < 180 - fA;
    p->isM = angle > 90;
    if (!p->isT && !p->isM) { p->sectorType = 0; p->newAngle = angle; }
    if (p->isT && !p->isM) { p->sectorType = 1; p->newAngle = 90 - angle; }
    if (p->isT && p->isM) { p->sectorType = 2; p->newAngle = angle - 90; }
    if (!p->isT && p->isM) { p->sectorType = 3; p->newAngle = 180 - angle; }
    p->scale = 1 / cos(torads * p->newAngle);
    p->skewness = tan(p->newAngle * torads);
    p->dim_i = p->isT ? dimy : dimx;
    p->dim_o = p->isT ? dimx : dimy;
    p->offset = (int)(p->skewness * (p->dim_i - 1));
    p->skewHeight = p->newAngle ? p->dim_o + p->offset + 1 : p->dim_o;
    p->skewWidth = p->dim_i;
    for (int j = 0; j < p->dim_i; j++) {
        double drift = p->skewness * j;
        target[j] = (int)drift;
        weight[j] = (float)(drift - target[j]);
    }
}

/* skewEngine.cuh:572-596 (skewShape), including the `j > last` comparison
 * that leaves some rows r < offset one cell short. */
void orc_skew_shape(const orc_params *p, const int *target, uint16_t *first, uint16_t *last)
{
    for (int i = 0; i < p->dim_o; i++) first[i] = 0;
    for (int i = p->offset; i < p->skewHeight; i++) last[i] = (uint16_t)p->dim_l;
    for (int i = 0; i < p->offset; i++) last[i] = 0;
    for (int i = p->dim_o; i < p->skewHeight; i++) first[i] = (uint16_t)p->dim_l;
    for (int j = 0; j < p->dim_i; j++) {
        int rows = target[j];
        if (j > last[rows]) last[rows] = (uint16_t)(j + 1);
        if (j > last[rows + 1]) last[rows + 1] = (uint16_t)(j + 1);
        int rowe = target[j] + p->dim_o - 1;
        if (j < first[rowe]) first[rowe] = (uint16_t)j;
        if (j < first[rowe + 1]) first[rowe + 1] = (uint16_t)j;
    }
}

/* skewEngine.cuh:962-987 (prepare): the three mirrored/transposed copies. */
void orc_prepare(const float *in0, int dimx, int dimy, float *in1, float *in2, float *in3)
{
    for (int i = 0; i < dimy; i++) {
        int ci = dimy - 1 - i;
        for (int j = 0; j < dimx; j++) {
            int cj = dimx - 1 - j;
            float val = in0[dimx * i + j];
            in1[dimy * cj + ci] = val;
            in2[dimy * j + ci] = val;
            in3[dimx * i + cj] = val;
        }
    }
}

/* skewEngine.cuh:541-570 (resetSkew, CPU branch) + :612-621 (skew, CPU
 * branch): two-tap scatter into the skewed layout. */
void orc_skew(const orc_params *p, const int *target, const float *weight, const float *src, float *S)
{
    int dim_i = p->dim_i;
    memset(S, 0, sizeof(float) * (size_t)p->skewHeight * dim_i);
    for (int i = 0; i < p->dim_o; i++) {
        for (int j = 0; j < dim_i; j++) {
            S[(size_t)(i + target[j]) * dim_i + j] += (1.0f - weight[j]) * src[(size_t)dim_i * i + j];
            S[(size_t)(i + target[j] + 1) * dim_i + j] += weight[j] * src[(size_t)dim_i * i + j];
        }
    }
}

static inline float stdmaxf(float a, float b) { return (a < b) ? b : a; } /* std::max */

/* skewEngine.cuh:755-793 (deskew, CPU branch).  isMax=0: weighted sum;
 * isMax=1: max then angle encoding in the low decimal digits (:782). */
void orc_deskew(const orc_params *p, const int *target, const float *weight, const float *SO, float *O, int isMax)
{
    int dim_i = p->dim_i;
    if (isMax == 1) {
        for (int i = 0; i < p->dim_o; i++)
            for (int j = 0; j < dim_i; j++) {
                size_t o = (size_t)i * dim_i + j;
                O[o] = stdmaxf(O[o], stdmaxf(SO[(size_t)(i + target[j]) * dim_i + j],
                                             SO[(size_t)(i + target[j] + 1) * dim_i + j]));
                O[o] = (float)(p->a + 1000 * (int)(O[o] / 1000));
            }
    } else {
        for (int i = 0; i < p->dim_o; i++)
            for (int j = 0; j < dim_i; j++) {
                size_t o = (size_t)i * dim_i + j;
                O[o] += (1.0f - weight[j]) * SO[(size_t)(i + target[j]) * dim_i + j] +
                        weight[j] * SO[(size_t)(i + target[j] + 1) * dim_i + j];
            }
    }
}

/* skewEngine.cuh:861-914 (reduce): inverse orientation maps, buffers added
 * (or maxed) in the order o0, o1, o2, o3; absent buffers (NULL) skipped. */
#define ORC_MAX(a, b) ((a) > (b) ? (a) : (b))
void orc_reduce(int dimx, int dimy, float *out, float *O0, float *O1, float *O2, float *O3, int isMax)
{
    for (int i = 0; i < dimy; i++) {
        int ci = dimy - 1 - i;
        for (int j = 0; j < dimx; j++) {
            int cj = dimx - 1 - j;
            size_t o = (size_t)i * dimx + j;
            if (isMax == 1) {
                if (O0) out[o] = ORC_MAX(out[o], O0[(size_t)i * dimx + j]);
                if (O1) out[o] = ORC_MAX(out[o], O1[(size_t)cj * dimy + ci]);
                if (O2) out[o] = ORC_MAX(out[o], O2[(size_t)j * dimy + ci]);
                if (O3) out[o] = ORC_MAX(out[o], O3[(size_t)i * dimx + cj]);
            } else {
                if (O0) out[o] += O0[(size_t)i * dimx + j];
                if (O1) out[o] += O1[(size_t)cj * dimy + ci];
                if (O2) out[o] += O2[(size_t)j * dimy + ci];
                if (O3) out[o] += O3[(size_t)i * dimx + cj];
            }
        }
    }
}

/* helper.h:59-106 (fft): in-place radix-2 decimation-in-frequency with
 * twiddles generated by repeated complex squaring/multiplication, then the
 * bit-reversal permutation, optional 1/sqrt(N).  Complex products are
 * (ac-bd, ad+bc) with each product rounded, as std::complex<double> does. */
void orc_fft(double *re, double *im, unsigned N, int normalize)
{
    unsigned k = N, n;
    double thetaT = 3.14159265358979323846264338328L / N;
    double pr = cos(thetaT), pi = -sin(thetaT);
    while (k > 1) {
        n = k;
        k >>= 1;
        {   /* phiT = phiT * phiT */
            double ac = pr * pr, bd = pi * pi, ad = pr * pi, bc = pi * pr;
            pr = ac - bd;
            pi = ad + bc;
        }
        double Tr = 1.0, Ti = 0.0;
        for (unsigned l = 0; l < k; l++) {
            for (unsigned a = l; a < N; a += n) {
                unsigned b = a + k;
                double tr = re[a] - re[b], ti = im[a] - im[b];
                re[a] += re[b];
                im[a] += im[b];
                double ac = tr * Tr, bd = ti * Ti, ad = tr * Ti, bc = ti * Tr;
                re[b] = ac - bd;
                im[b] = ad + bc;
            }
            /* T *= phiT */
            double ac = Tr * pr, bd = Ti * pi, ad = Tr * pi, bc = Ti * pr;
            Tr = ac - bd;
            Ti = ad + bc;
        }
    }
    unsigned m = (unsigned)log2(N);
    for (unsigned a = 0; a < N; a++) {
        unsigned b = a;
        b = (((b & 0xaaaaaaaa) >> 1) | ((b & 0x55555555) << 1));
        b = (((b & 0xcccccccc) >> 2) | ((b & 0x33333333) << 2));
        b = (((b & 0xf0f0f0f0) >> 4) | ((b & 0x0f0f0f0f) << 4));
        b = (((b & 0xff00ff00) >> 8) | ((b & 0x00ff00ff) << 8));
        b = ((b >> 16) | (b << 16)) >> (32 - m);
        if (b > a) {
            double t = re[a]; re[a] = re[b]; re[b] = t;
            t = im[a]; im[a] = im[b]; im[b] = t;
        }
    }
    if (!normalize) return;
    double f = 1.0 / sqrt(N);
    for (unsigned i = 0; i < N; i++) {
        /* x[i] *= Complex(f, 0): (re*f - im*0, re*0 + im*f) */
        double r = re[i] * f - im[i] * 0.0;
        double q = re[i] * 0.0 + im[i] * f;
        re[i] = r;
        im[i] = q;
    }
}

/* helper.h:108-114 (ifft): conj, fft, conj, divide by N. */
void orc_ifft(double *re, double *im, unsigned N)
{
    for (unsigned i = 0; i < N; i++) im[i] = -im[i];
    orc_fft(re, im, N, 0);
    for (unsigned i = 0; i < N; i++) {
        im[i] = -im[i];
        re[i] = re[i] / N;
        im[i] = im[i] / N;
    }
}

/* kernelSDEM.h:189-217, one observer: right scan then left scan, cv kept
 * across both; a visible run still open at the row end is not counted. */
float orc_sdem_observer(const float *row, int inicio, int final, int j, float povh,
                        uint8_t *vis_right, uint8_t *vis_left)
{
    float cv = 0, max_angle, open_delta_d, delta_d, cAngle;
    int visible, above, opening, closing;
    float h = row[j] + povh;
    max_angle = -2000;
    visible = 1;
    open_delta_d = 0;
    for (int k = j + 1; k >= inicio && k < final; k++) {
        delta_d = (float)(k - j);
        cAngle = (row[k] - h) / delta_d;
        above = (cAngle > max_angle);
        opening = above && (!visible);
        closing = (!above) && (visible);
        visible = above;
        max_angle = stdmaxf(cAngle, max_angle);
        if (opening) open_delta_d = delta_d;
        if (closing) cv += (delta_d * delta_d - open_delta_d * open_delta_d);
        if (vis_right) vis_right[k - j - 1] = (uint8_t)above;
    }
    max_angle = -2000;
    visible = 1;
    open_delta_d = 0;
    for (int k = j - 1; k >= inicio && k < final; k--) {
        delta_d = (float)(j - k);
        cAngle = (row[k] - h) / delta_d;
        above = (cAngle > max_angle);
        opening = above && (!visible);
        closing = (!above) && (visible);
        visible = above;
        max_angle = stdmaxf(cAngle, max_angle);
        if (opening) open_delta_d = delta_d;
        if (closing) cv += (delta_d * delta_d - open_delta_d * open_delta_d);
        if (vis_left) vis_left[j - k - 1] = (uint8_t)above;
    }
    return cv;
}

/* kernelSDEM.h:168-222 (kernelV3).  r = (float)pow(scale,2): the compiler
 * folds pow(x,2) into x*x. */
void orc_op_sdem(const orc_params *p, const uint16_t *first, const uint16_t *last,
                 const float *S, float *SO, float povh)
{
    float r = (float)(p->scale * p->scale);
    for (int i = 0; i < p->skewHeight; i++) {
        int inicio = first[i], final = last[i];
        const float *row = S + (size_t)p->skewWidth * i;
        for (int j = inicio; j < final; j++)
            SO[(size_t)p->skewWidth * i + j] = orc_sdem_observer(row, inicio, final, j, povh, NULL, NULL) * r;
    }
}

/* kernelUNIT.h:16-23 (identity). */
void orc_op_identity(const orc_params *p, const uint16_t *first, const uint16_t *last,
                     const float *S, float *SO)
{
    for (int i = 0; i < p->skewHeight; i++) {
        size_t k = (size_t)p->skewWidth * i;
        for (int j = first[i]; j < last[i]; j++) SO[k + j] = S[k + j];
    }
}

/* kernelRADN.h:221-258 (radon, global mode LRT=false): row sums in double,
 * FFT of length H2 = 2^ceil(log2 H), one-sided ramp k/H for k <= H/2,
 * IFFT, real part smeared back along the row. */
void orc_op_radon(const orc_params *p, const uint16_t *first, const uint16_t *last,
                  const float *S, float *SO)
{
    int H = p->skewHeight;
    int H2 = (int)pow(2, ceil(log2(H)));
    double *re = (double *)calloc((size_t)H2, sizeof(double));
    double *im = (double *)calloc((size_t)H2, sizeof(double));
    for (int i = 0; i < H; i++) {
        size_t k = (size_t)p->skewWidth * i;
        for (int j = first[i]; j < last[i]; j++) re[i] = re[i] + S[k + j];
    }
    orc_fft(re, im, (unsigned)H2, 0);
    for (int i = 0; i < H2; i++) {
        if (i > H / 2) {
            re[i] = 0;
            im[i] = 0;
        } else {
            double g = (double)i / H;
            re[i] = g * re[i];
            im[i] = g * im[i];
        }
    }
    orc_ifft(re, im, (unsigned)H2);
    for (int i = 0; i < H; i++) {
        size_t k = (size_t)p->skewWidth * i;
        for (int j = first[i]; j < last[i]; j++) SO[k + j] = (float)(SO[k + j] + re[i]);
    }
    free(re);
    free(im);
}

/* kernelBLUR.h:88-110 (getBlurParameter2). */
double orc_blur_parameter2(const double *v, int n, int *imax_out)
{
    double val = 0;
    int imax = 0, i;
    for (i = 5; i < n; i++)
        if (v[i] > v[i - 1]) break;
    if (i == n) {
        if (imax_out) *imax_out = 0;
        return 0;
    }
    val = v[i - 1];
    imax = i++;
    for (; i < n; i++)
        if (v[i] > val) {
            val = v[i];
            imax = i;
        }
    if (imax == 0) val = 0;
    if (imax_out) *imax_out = imax;
    return val;
}

#define ORC_WIN 64 /* winSize from kernelRADN.h:19 (included first, main.cu:64-65) */

/* kernelBLUR.h:153-228 (lineCepstrum) with select_subarray (:69-85):
 * 64-tap window (zero outside [first,last)), FFT/8, log power of the first
 * 32 bins, FFT/sqrt(32), v_k = -min(0, Re C_k) for k < 16, peak rule,
 * times 10^4. */
float orc_line_cepstrum_ex(int j, const float *row, int first, int last, double *logps_out, double *ceps_out)
{
    double re[ORC_WIN], im[ORC_WIN];
    int half = ORC_WIN / 2;
    for (int m = 0; m < ORC_WIN; m++) {
        int index = j - half + m;
        re[m] = (index < first || index >= last) ? 0.0 : (double)row[index];
        im[m] = 0.0;
    }
    orc_fft(re, im, ORC_WIN, 1);
    double re2[ORC_WIN / 2], im2[ORC_WIN / 2];
    for (int k = 0; k < ORC_WIN / 2; k++) {
        double PS = re[k] * re[k] + im[k] * im[k];
        re2[k] = log(PS);
        im2[k] = 0.0;
        if (logps_out) logps_out[k] = re2[k];
    }
    orc_fft(re2, im2, ORC_WIN / 2, 1);
    double samples[ORC_WIN / 4];
    for (int k = 0; k < ORC_WIN / 4; k++) {
        double x = re2[k];
        samples[k] = -((x < 0.0) ? x : 0.0); /* -std::min(0.0, x) */
        if (ceps_out) ceps_out[k] = x;
    }
    double val = orc_blur_parameter2(samples, ORC_WIN / 4, NULL);
    return (float)(val * 10000);
}

float orc_line_cepstrum(int j, const float *row, int first, int last)
{
    return orc_line_cepstrum_ex(j, row, first, last, NULL, NULL);
}

/* kernelRADN.h:114-182 (pointRadon, local Radon transform, LRT=true) for
 * skewed cell (i, j): 64 rows k = i-32..i+31 of 65-column window sums
 * (columns j-32..j+32 inside [first[k], last[k]), zero for rows outside
 * [0, H)) times a Gaussian of sigma = (64/5)/scale, FFT/8, a sine ramp on
 * the real parts of bins 0..25 (zero above 25.6), FFT/8 again, and the
 * squared magnitude at the first rise after bin 5 (getParameter :96-108). */
/* samples_out (4 x 64 floats, may be NULL): the vectors pointRadon records
 * for a sample point (kernelRADN.h:164-170): the windowed sums (input1), the
 * Gaussian-weighted row through the cell (axis, with the extents of row k of
 * the window as the reference reads them; rows outside [0, H) give 0), the
 * filtered spectrum (input2) and the real part of its transform (data2). */
static double point_radon_core(const orc_params *p, const uint16_t *first, const uint16_t *last, const float *S,
                               int i, int j, double *spec_out, float *samples_out)
{
    const int w = p->skewWidth, h = p->skewHeight, w2 = ORC_WIN / 2;
    const double dws = ORC_WIN;
    double sigma = ORC_WIN / 5.0;
    sigma = sigma / p->scale;
    double re[ORC_WIN], im[ORC_WIN], axis[ORC_WIN];
    for (int k = i - w2; k < i + w2; k++) {
        double accum = 0;
        int fi = 0, la = 0;
        if (k >= 0 && k < h) {
            fi = first[k];
            la = last[k];
            for (int l = j - w2; l <= j + w2; l++)
                accum += (l < fi || l >= la) ? 0 : S[(size_t)k * w + l];
        }
        int kk = k - (i - w2);
        double gauss = exp(-(kk - w2) * (kk - w2) / (2 * sigma * sigma));
        re[kk] = accum * gauss;
        im[kk] = 0.0;
        int c = j + k - i;
        axis[kk] = (c < fi) ? 0 : (c >= la ? 0 : gauss * S[(size_t)i * w + c]);
    }
    if (samples_out)
        for (int k = 0; k < ORC_WIN; k++) {
            samples_out[k] = (float)re[k];
            samples_out[ORC_WIN + k] = (float)axis[k];
        }
    orc_fft(re, im, ORC_WIN, 1);
    double cutoff = 0.4;
    double re2[ORC_WIN], im2[ORC_WIN];
    for (int k = 0; k < ORC_WIN; k++) {
        re2[k] = k > cutoff * dws ? 0 : sin((k / (cutoff * dws - 1)) * (M_PI / 2)) * re[k];
        im2[k] = 0.0;
    }
    if (samples_out)
        for (int k = 0; k < ORC_WIN; k++) samples_out[2 * ORC_WIN + k] = (float)re2[k];
    orc_fft(re2, im2, ORC_WIN, 1);
    if (samples_out)
        for (int k = 0; k < ORC_WIN; k++) samples_out[3 * ORC_WIN + k] = (float)re2[k];
    if (spec_out)
        for (int k = 0; k < ORC_WIN; k++) spec_out[k] = re2[k] * re2[k] + im2[k] * im2[k];
    int n = w2, ii;
    for (ii = 5; ii < n; ii++)
        if (re2[ii] * re2[ii] + im2[ii] * im2[ii] > re2[ii - 1] * re2[ii - 1] + im2[ii - 1] * im2[ii - 1]) break;
    if (ii == n) return 0.0;
    return re2[ii - 1] * re2[ii - 1] + im2[ii - 1] * im2[ii - 1];
}

double orc_point_radon(const orc_params *p, const uint16_t *first, const uint16_t *last, const float *S,
                       int i, int j, double *spec_out)
{
    return point_radon_core(p, first, last, S, i, j, spec_out, NULL);
}

double orc_point_radon_samples(const orc_params *p, const uint16_t *first, const uint16_t *last, const float *S,
                               int i, int j, float *samples_out)
{
    return point_radon_core(p, first, last, S, i, j, NULL, samples_out);
}

/* kernelRADN.h:214-220 (radon with LRT=true). */
void orc_op_radon_local(const orc_params *p, const uint16_t *first, const uint16_t *last,
                        const float *S, float *SO)
{
    for (int i = 0; i < p->skewHeight; i++) {
        size_t k = (size_t)p->skewWidth * i;
        for (int j = first[i]; j < last[i]; j++) SO[k + j] = (float)orc_point_radon(p, first, last, S, i, j, NULL);
    }
}

/* kernelBLUR.h:267-302 (cepstrum). */
void orc_op_cepstrum(const orc_params *p, const uint16_t *first, const uint16_t *last,
                     const float *S, float *SO)
{
    for (int i = 0; i < p->skewHeight; i++) {
        size_t k = (size_t)p->skewWidth * i;
        for (int j = first[i]; j < last[i]; j++)
            SO[k + j] = orc_line_cepstrum(j, S + k, first[i], last[i]);
    }
}

/* main.h:94-101: largest divisor of 180 not above the processor count. */
int orc_default_threads(void)
{
    int n = omp_get_num_procs();
    while (180 % n) n--;
    return n;
}

/* One engine per thread: the fields of skewEngine<float> that the CPU path
 * touches (skewEngine.cuh:253-367, buffers from skewAlloc :418-426). */
typedef struct {
    int dimx, dimy, N;
    const float *in[4];
    float *out[4];
    int *target;
    float *weight;
    uint16_t *first, *last;
    float *S, *SO;
    orc_params p;
} orc_engine;

static int eng_init(orc_engine *e, int dimx, int dimy, const float *in4[4])
{
    memset(e, 0, sizeof(*e));
    e->dimx = dimx;
    e->dimy = dimy;
    e->N = dimx * dimy;
    for (int t = 0; t < 4; t++) e->in[t] = in4[t];
    int dim_l = dimx > dimy ? dimx : dimy;
    size_t sk = (size_t)e->N * 2 + dim_l;
    e->target = (int *)calloc((size_t)dim_l, sizeof(int));
    e->weight = (float *)calloc((size_t)dim_l, sizeof(float));
    e->first = (uint16_t *)calloc((size_t)2 * dim_l + 1, sizeof(uint16_t));
    e->last = (uint16_t *)calloc((size_t)2 * dim_l + 1, sizeof(uint16_t));
    /* zero-filled so that the rows past skewHeight that deskew can touch
     * when newAngle == 0 start out deterministic */
    e->S = (float *)calloc(sk, sizeof(float));
    e->SO = (float *)calloc(sk, sizeof(float));
    return (e->target && e->weight && e->first && e->last && e->S && e->SO) ? 0 : -1;
}

static void eng_free(orc_engine *e)
{
    free(e->target); free(e->weight); free(e->first); free(e->last);
    free(e->S); free(e->SO);
    for (int t = 0; t < 4; t++) free(e->out[t]);
}

/* skew(angle) :605-621 then kernel() :694-696 then deskew(isMax) :755-793. */
static void eng_sector(orc_engine *e, int alg, int angle, float povh, int do_deskew)
{
    orc_params *p = &e->p;
    orc_compute_params(e->dimx, e->dimy, angle, p, e->target, e->weight);
    orc_skew_shape(p, e->target, e->first, e->last);
    /* resetSkew zeroes skewHeight*dim_i of both buffers */
    memset(e->SO, 0, sizeof(float) * (size_t)p->skewHeight * p->dim_i);
    orc_skew(p, e->target, e->weight, e->in[p->sectorType], e->S);
    switch (alg) {
    case ORC_SDEM: orc_op_sdem(p, e->first, e->last, e->S, e->SO, povh); break;
    case ORC_RADON: orc_op_radon(p, e->first, e->last, e->S, e->SO); break;
    case ORC_RADON_LOCAL: orc_op_radon_local(p, e->first, e->last, e->S, e->SO); break;
    case ORC_CEPSTRUM: orc_op_cepstrum(p, e->first, e->last, e->S, e->SO); break;
    default: orc_op_identity(p, e->first, e->last, e->S, e->SO); break;
    }
    if (!do_deskew) return;
    int t = p->sectorType;
    if (!e->out[t]) e->out[t] = (float *)calloc((size_t)e->N, sizeof(float));
    orc_deskew(p, e->target, e->weight, e->SO, e->out[t], alg == ORC_CEPSTRUM ? 1 : 0);
}

int orc_execute(int alg, int dimx, int dimy, const float *input, float *out,
                int nthreads, int s0, int s1, float povh, double *seconds)
{
    size_t N = (size_t)dimx * dimy;
    float *in1 = (float *)malloc(N * sizeof(float));
    float *in2 = (float *)malloc(N * sizeof(float));
    float *in3 = (float *)malloc(N * sizeof(float));
    if (!in1 || !in2 || !in3) return -1;
    orc_prepare(input, dimx, dimy, in1, in2, in3);
    const float *in4[4] = {input, in1, in2, in3};
    memset(out, 0, N * sizeof(float));
    if (nthreads <= 0) nthreads = orc_default_threads();
    int isMax = alg == ORC_CEPSTRUM ? 1 : 0;
    int err = 0;
    double t1 = omp_get_wtime();
#pragma omp parallel num_threads(nthreads)
    {
        orc_engine e;
        if (eng_init(&e, dimx, dimy, in4) != 0) {
#pragma omp atomic write
            err = -1;
        }
#pragma omp barrier
#pragma omp for schedule(dynamic) nowait
        for (int i = s0; i < s1; i++)
            if (!err) eng_sector(&e, alg, i, povh, 1);
#pragma omp critical
        {
            if (!err) orc_reduce(dimx, dimy, out, e.out[0], e.out[1], e.out[2], e.out[3], isMax);
        }
        eng_free(&e);
    }
    double t2 = omp_get_wtime();
    if (seconds) *seconds = t2 - t1;
    free(in1); free(in2); free(in3);
    return err;
}

/* FP64 restatement of the total-viewshed farm (kernelV3, kernelSDEM.h:168-222,
 * driven as main.cu:71-172): the same index maps and scans, with the skew
 * weights (skewness*j - target[j]), the skewed heights, the elevation angles,
 * the observer height, cv, the area factor scale^2, the deskew and the reduce
 * all in double.  Its totals are what the FP32 algorithm approximates: the
 * gap measures the float rounding of the reference's own arithmetic. */
static double observer_f64(const double *row, int inicio, int final, int j, double povh)
{
    double cv = 0, max_angle, open_delta_d, delta_d, cAngle;
    int visible, above;
    double h = row[j] + povh;
    for (int dir = 1; dir >= -1; dir -= 2) {
        max_angle = -2000;
        visible = 1;
        open_delta_d = 0;
        for (int k = j + dir; k >= inicio && k < final; k += dir) {
            delta_d = (double)(dir > 0 ? k - j : j - k);
            cAngle = (row[k] - h) / delta_d;
            above = cAngle > max_angle;
            if (above && !visible) open_delta_d = delta_d;
            if (!above && visible) cv += delta_d * delta_d - open_delta_d * open_delta_d;
            visible = above;
            if (cAngle > max_angle) max_angle = cAngle;
        }
    }
    return cv;
}

int orc_execute_sdem_f64(int dimx, int dimy, const float *input, double *out, int nthreads, int s0, int s1,
                         double povh)
{
    size_t N = (size_t)dimx * dimy;
    float *in1 = (float *)malloc(N * sizeof(float));
    float *in2 = (float *)malloc(N * sizeof(float));
    float *in3 = (float *)malloc(N * sizeof(float));
    if (!in1 || !in2 || !in3) return -1;
    orc_prepare(input, dimx, dimy, in1, in2, in3);
    const float *in4[4] = {input, in1, in2, in3};
    memset(out, 0, N * sizeof(double));
    if (nthreads <= 0) nthreads = orc_default_threads();
    int dim_l = dimx > dimy ? dimx : dimy;
    int err = 0;
#pragma omp parallel num_threads(nthreads)
    {
        int *target = (int *)calloc((size_t)dim_l, sizeof(int));
        float *weight = (float *)calloc((size_t)dim_l, sizeof(float));
        double *wd = (double *)calloc((size_t)dim_l, sizeof(double));
        uint16_t *first = (uint16_t *)calloc(2 * (size_t)dim_l + 1, sizeof(uint16_t));
        uint16_t *last = (uint16_t *)calloc(2 * (size_t)dim_l + 1, sizeof(uint16_t));
        double *S = (double *)malloc((2 * N + dim_l) * sizeof(double));
        double *SO = (double *)malloc((2 * N + dim_l) * sizeof(double));
        double *acc[4] = {NULL, NULL, NULL, NULL};
        if (!target || !weight || !wd || !first || !last || !S || !SO) {
#pragma omp atomic write
            err = -1;
        }
#pragma omp barrier
#pragma omp for schedule(dynamic) nowait
        for (int a = s0; a < s1; a++) {
            if (err) continue;
            orc_params p;
            orc_compute_params(dimx, dimy, a, &p, target, weight);
            orc_skew_shape(&p, target, first, last);
            int di = p.dim_i, H = p.skewHeight;
            for (int j = 0; j < di; j++) wd[j] = p.skewness * j - target[j];
            memset(S, 0, (size_t)H * di * sizeof(double));
            memset(SO, 0, (size_t)H * di * sizeof(double));
            const float *src = in4[p.sectorType];
            for (int i = 0; i < p.dim_o; i++)
                for (int j = 0; j < di; j++) {
                    double v = src[(size_t)di * i + j];
                    S[(size_t)(i + target[j]) * di + j] += (1.0 - wd[j]) * v;
                    S[(size_t)(i + target[j] + 1) * di + j] += wd[j] * v;
                }
            double r = p.scale * p.scale;
            for (int i = 0; i < H; i++)
                for (int j = first[i]; j < last[i]; j++)
                    SO[(size_t)di * i + j] = observer_f64(S + (size_t)di * i, first[i], last[i], j, povh) * r;
            int t = p.sectorType;
            if (!acc[t]) acc[t] = (double *)calloc(N, sizeof(double));
            if (!acc[t]) {
#pragma omp atomic write
                err = -1;
                continue;
            }
            for (int i = 0; i < p.dim_o; i++)
                for (int j = 0; j < di; j++)
                    acc[t][(size_t)i * di + j] += (1.0 - wd[j]) * SO[(size_t)(i + target[j]) * di + j] +
                                                  wd[j] * SO[(size_t)(i + target[j] + 1) * di + j];
        }
#pragma omp critical
        {
            if (!err)
                for (int i = 0; i < dimy; i++) {
                    int ci = dimy - 1 - i;
                    for (int j = 0; j < dimx; j++) {
                        int cj = dimx - 1 - j;
                        double *o = out + (size_t)i * dimx + j;
                        if (acc[0]) *o += acc[0][(size_t)i * dimx + j];
                        if (acc[1]) *o += acc[1][(size_t)cj * dimy + ci];
                        if (acc[2]) *o += acc[2][(size_t)j * dimy + ci];
                        if (acc[3]) *o += acc[3][(size_t)i * dimx + cj];
                    }
                }
        }
        for (int t = 0; t < 4; t++) free(acc[t]);
        free(target); free(weight); free(wd); free(first); free(last); free(S); free(SO);
    }
    free(in1); free(in2); free(in3);
    return err;
}

int orc_sector(int alg, int dimx, int dimy, const float *input, int angle, float povh,
               orc_params *p, int *target, float *weight, uint16_t *first, uint16_t *last,
               float *S, float *SO)
{
    size_t N = (size_t)dimx * dimy;
    float *in1 = (float *)malloc(N * sizeof(float));
    float *in2 = (float *)malloc(N * sizeof(float));
    float *in3 = (float *)malloc(N * sizeof(float));
    if (!in1 || !in2 || !in3) return -1;
    orc_prepare(input, dimx, dimy, in1, in2, in3);
    const float *in4[4] = {input, in1, in2, in3};
    orc_engine e;
    if (eng_init(&e, dimx, dimy, in4) != 0) return -1;
    eng_sector(&e, alg, angle, povh, 0);
    int dim_l = dimx > dimy ? dimx : dimy;
    size_t sk = N * 2 + dim_l;
    *p = e.p;
    memcpy(target, e.target, sizeof(int) * dim_l);
    memcpy(weight, e.weight, sizeof(float) * dim_l);
    memcpy(first, e.first, sizeof(uint16_t) * (2 * dim_l + 1));
    memcpy(last, e.last, sizeof(uint16_t) * (2 * dim_l + 1));
    memcpy(S, e.S, sizeof(float) * sk);
    memcpy(SO, e.SO, sizeof(float) * sk);
    eng_free(&e);
    free(in1); free(in2); free(in3);
    return 0;
}

void orc_work(int dimx, int dimy, int s0, int s1, double *cells, double *steps)
{
    int dim_l = dimx > dimy ? dimx : dimy;
    int *target = (int *)calloc((size_t)dim_l, sizeof(int));
    float *weight = (float *)calloc((size_t)dim_l, sizeof(float));
    uint16_t *first = (uint16_t *)calloc((size_t)2 * dim_l + 1, sizeof(uint16_t));
    uint16_t *last = (uint16_t *)calloc((size_t)2 * dim_l + 1, sizeof(uint16_t));
    double c = 0, s = 0;
    for (int a = s0; a < s1; a++) {
        orc_params p;
        orc_compute_params(dimx, dimy, a, &p, target, weight);
        orc_skew_shape(&p, target, first, last);
        for (int r = 0; r < p.skewHeight; r++) {
            double L = (double)last[r] - (double)first[r];
            if (L > 0) {
                c += L;
                s += L * (L - 1);
            }
        }
    }
    free(target); free(weight); free(first); free(last);
    if (cells) *cells = c;
    if (steps) *steps = s;
}
