/*
 * ske_oracle.h -- CPU restatement of the reference skewEngine CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (skewengine_amd) never does.
 *
 * Every function restates one piece of the reference and cites it as
 * file:line relative to the reference repository root.
 *
 * Parity pinning (what checks each function beyond re-typing it) is enforced
 * by tests/test_oracle_*.py.
 */
#ifndef SKE_ORACLE_H
#define SKE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-sector parameters: the public fields set by computeParams
 * (src/skewEngine.cuh:503-539). */
typedef struct orc_params {
    int a;            /* sector angle, 0..179 */
    int newAngle;     /* folded angle 0..45 for square inputs */
    int sectorType;   /* 0..3 */
    int isT, isM;
    int dim_i, dim_o, dim_l;
    int offset, skewHeight, skewWidth;
    double scale;     /* 1/cos(newAngle) */
    double skewness;  /* tan(newAngle) */
    float fAngle;     /* atan2(dimy,dimx) in degrees, stored as float */
} orc_params;

enum { ORC_SDEM = 0, ORC_IDENTITYDEM = 1, ORC_RADON = 2, ORC_IDENTITYIMG = 3, ORC_CEPSTRUM = 4 };
/* Oracle-only id: the Radon operator with the reference's LRT flag set
 * (local transform, kernelRADN.h:212-220). */
enum { ORC_RADON_LOCAL = 5 };

float orc_fangle(int dimx, int dimy);
void orc_compute_params(int dimx, int dimy, int angle, orc_params *p, int *target, float *weight);
void orc_skew_shape(const orc_params *p, const int *target, uint16_t *first, uint16_t *last);
void orc_prepare(const float *in0, int dimx, int dimy, float *in1, float *in2, float *in3);
void orc_skew(const orc_params *p, const int *target, const float *weight, const float *src, float *S);
void orc_deskew(const orc_params *p, const int *target, const float *weight, const float *SO, float *O, int isMax);
void orc_reduce(int dimx, int dimy, float *out, float *O0, float *O1, float *O2, float *O3, int isMax);

void orc_fft(double *re, double *im, unsigned N, int normalize);
void orc_ifft(double *re, double *im, unsigned N);

void orc_op_sdem(const orc_params *p, const uint16_t *first, const uint16_t *last,
                 const float *S, float *SO, float povh);
void orc_op_identity(const orc_params *p, const uint16_t *first, const uint16_t *last,
                     const float *S, float *SO);
void orc_op_radon(const orc_params *p, const uint16_t *first, const uint16_t *last,
                  const float *S, float *SO);
void orc_op_cepstrum(const orc_params *p, const uint16_t *first, const uint16_t *last,
                     const float *S, float *SO);
void orc_op_radon_local(const orc_params *p, const uint16_t *first, const uint16_t *last,
                        const float *S, float *SO);
/* One cell of the local transform; spec_out (64 doubles, may be NULL)
 * receives the squared magnitudes the peak rule scans. */
double orc_point_radon(const orc_params *p, const uint16_t *first, const uint16_t *last, const float *S,
                       int i, int j, double *spec_out);
/* The same cell, returning the four 64-float vectors the reference records
 * for a sample point (kernelRADN.h:164-170): windowed sums, axis, filtered
 * spectrum, real part of its transform. */
double orc_point_radon_samples(const orc_params *p, const uint16_t *first, const uint16_t *last, const float *S,
                               int i, int j, float *samples_out);
double orc_blur_parameter2(const double *v, int n, int *imax);
float orc_line_cepstrum(int j, const float *row, int first, int last);
/* The same, also returning the intermediates: the 32 log powers ln|X_k|^2
 * and the 16 cepstrum values Re C_k whose -min(0, .) the peak rule scans
 * (either pointer may be NULL). */
float orc_line_cepstrum_ex(int j, const float *row, int first, int last, double *logps_out, double *ceps_out);

/* One observer's two line-of-sight scans on a staged row; writes the
 * visibility decision of every step (1 = above the horizon) and returns cv. */
float orc_sdem_observer(const float *row, int first, int last, int j, float povh,
                        uint8_t *vis_right, uint8_t *vis_left);

int orc_default_threads(void);

/* The OpenMP sector farm of main.cu:71-172 over sectors [s0, s1).  Includes
 * allocate/prepare (untimed, as main.cu:186) and times the farm like
 * main.cu:187-192.  Returns 0 on success. */
int orc_execute(int alg, int dimx, int dimy, const float *input, float *out,
                int nthreads, int s0, int s1, float povh, double *seconds);

/* FP64 variant of the total-viewshed farm over sectors [s0, s1): same maps
 * and algorithm, every value in double (out: dimx*dimy doubles). */
int orc_execute_sdem_f64(int dimx, int dimy, const float *input, double *out, int nthreads, int s0, int s1,
                         double povh);

/* Run one sector through skew -> operator on a fresh engine and return its
 * tables and skewed buffers (sized 2*N + dim_l floats each).  For tests. */
int orc_sector(int alg, int dimx, int dimy, const float *input, int angle, float povh,
               orc_params *p, int *target, float *weight, uint16_t *first, uint16_t *last,
               float *S, float *SO);

/* Work counts from the exact index maps: cells = sum L_r, steps = sum L_r(L_r-1)
 * over sectors [s0, s1). */
void orc_work(int dimx, int dimy, int s0, int s1, double *cells, double *steps);

#ifdef __cplusplus
}
#endif
#endif
