// planner.h -- host-side sector planning: the index maps of the reference's
// computeParams/skewShape (src/skewEngine.cuh:503-596), packed into the
// device tables the kernels read, plus the constant tables (exact-division
// reciprocals, FFT twiddles, cepstrum cosines).
#pragma once
#include <stdint.h>

#include <vector>

#include "ske_types.h"

namespace ske {

struct SectorPlan {
    int angle = 0, newAngle = 0, type = 0, isT = 0, isM = 0;
    int dim_i = 0, dim_o = 0, dim_l = 0, offset = 0, H = 0, skewWidth = 0;
    double scale = 1.0, skewness = 0.0;
    std::vector<int> target;      // dim_l entries (first dim_i meaningful)
    std::vector<float> weight;    // dim_l entries
    std::vector<uint16_t> first;  // 2*dim_l+1 entries, rows < H meaningful
    std::vector<uint16_t> last;
};

// fAngle = atan2(dimy,dimx)*180/M_PI as float (src/skewEngine.cuh:465).
float fangle(int dimx, int dimy);

// computeParams(angle) + skewShape() for one sector.
void plan_sector(int dimx, int dimy, int angle, SectorPlan& sp);

// A run of consecutive sectors of one image, laid out for the device.
struct ImagePlan {
    int dimx = 0, dimy = 0, s0 = 0, s1 = 0;
    std::vector<SecDesc> secs;
    std::vector<int> target;
    std::vector<float> weight;
    std::vector<uint32_t> fl;
    std::vector<RowWork> work;  // rows with at least one valid cell
    int maxL = 0;
    int maxH = 0;
    long long so_total = 0;     // arena floats (sum of H*dim_i)
    long long q_total = 0;      // sum of H
    int range[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // per-type [begin, end) into secs
    double cells = 0, steps = 0;              // sum L_r, sum L_r(L_r-1)
};

enum WorkOrder { kOrderNatural = 0, kOrderCostDesc = 1 };

// Plan sectors [s0, s1) of a dimx x dimy image.  `img` is the element offset
// of the image's four oriented copies; `so_base`/`q_base` offset its arena
// and projection storage.  Rows are sorted by descending L^2 cost when
// `order` is kOrderCostDesc (viewshed), otherwise image by image and by
// descending L within an image.
void plan_image(int dimx, int dimy, int s0, int s1, long long img, long long so_base, long long q_base,
                int order, ImagePlan& ip);

// Append sectors [s0, s1) of one more image to `ip` (arena and projection
// offsets continue from ip.so_total / ip.q_total); range8 receives the
// image's per-type [begin, end) indices into ip.secs.  Call plan_finish once
// after the last append to order the row work list.
void plan_append(int dimx, int dimy, int s0, int s1, long long img, ImagePlan& ip, int range8[8]);
void plan_finish(ImagePlan& ip, int order);

// {RN(1/d), RN(1/d - RN(1/d)), d, d*d} for d = 0..n-1, followed by the
// pairs {RN(1/d), RN(1/d - RN(1/d))} for d = 0..n+31 (the viewshed's
// block-preloaded table).
void recip_table(int n, std::vector<float>& tab4);

// Twiddles of the reference FFT (src/helper.h:63-81) for N = 2^m, m = 1..mmax,
// concatenated as (re, im) pairs; off[m] is the pair offset of N = 2^m.
// Within one N, stage n (n = N, N/2, ..., 2) occupies pairs [N-n, N-n+n/2).
void fft_twiddles(int mmax, std::vector<double>& tw, std::vector<int>& off);

// Cepstrum/local-Radon constants (doubles): [0,128) cos/sin(2 pi t/64)
// interleaved, [128,160) cos(2 pi t/32), [160,186) the local Radon's sine
// ramp sin(k/24.6 * pi/2), then from kLogTab (ske_types.h) the natural-log
// table of the cepstrum kernel (256 pairs c_i, -log c_i, then ln 2 as a
// double-double), then from kLrtC and kLrtY the local Radon's DFT matrices
// and from kBlurD the cepstrum's window DFT matrix.
void cepstrum_table(std::vector<double>& tab);

// Gaussian window of the local Radon transform, 64 weights per folded angle
// newAngle = 0..90.
void lrt_gauss_table(std::vector<double>& tab);

}  // namespace ske
