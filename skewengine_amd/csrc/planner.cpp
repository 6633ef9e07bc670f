// planner.cpp -- host-side sector planning (see planner.h).
//
// The maps depend on libm tan/cos in double, so they are computed here on the
// host, exactly as src/skewEngine.cuh:503-596 computes them, and uploaded.
#include "planner.h"

#include <math.h>

#include <algorithm>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace ske {

float fangle(int dimx, int dimy) { return (float)(atan2((double)dimy, (double)dimx) * 180 / M_PI); }

void plan_sector(int dimx, int dimy, int angle, SectorPlan& sp) {
    const double torads = M_PI / 180.0;  // src/skewEngine.cuh:258
    const int fA = (int)fangle(dimx, dimy);
    sp.dim_l = std::max(dimx, dimy);
    sp.angle = angle;
    // sector type and folded angle (:507-512)
    sp.isT = angle > fA && angle < 180 - fA;
    sp.isM = angle > 90;
    if (!sp.isT && !sp.isM) { sp.type = 0; sp.newAngle = angle; }
    if (sp.isT && !sp.isM) { sp.type = 1; sp.newAngle = 90 - angle; }
    if (sp.isT && sp.isM) { sp.type = 2; sp.newAngle = angle - 90; }
    if (!sp.isT && sp.isM) { sp.type = 3; sp.newAngle = 180 - angle; }
    sp.scale = 1 / cos(torads * sp.newAngle);      // :516
    sp.skewness = tan(sp.newAngle * torads);       // :522
    sp.dim_i = sp.isT ? dimy : dimx;               // :525-526
    sp.dim_o = sp.isT ? dimx : dimy;
    sp.offset = (int)(sp.skewness * (sp.dim_i - 1));             // :528
    sp.H = sp.newAngle ? sp.dim_o + sp.offset + 1 : sp.dim_o;    // :529
    sp.skewWidth = sp.dim_i;
    sp.target.assign(sp.dim_l, 0);
    sp.weight.assign(sp.dim_l, 0.0f);
    for (int j = 0; j < sp.dim_i; j++) {                          // :533-537
        const double drift = sp.skewness * j;
        sp.target[j] = (int)drift;
        sp.weight[j] = (float)(drift - sp.target[j]);
    }
    // skewShape (:580-594).  The update compares j > last[] (not j+1 > last[]).
    sp.first.assign(2 * sp.dim_l + 1, 0);
    sp.last.assign(2 * sp.dim_l + 1, 0);
    for (int i = 0; i < sp.dim_o; i++) sp.first[i] = 0;
    for (int i = sp.offset; i < sp.H; i++) sp.last[i] = (uint16_t)sp.dim_l;
    for (int i = 0; i < sp.offset; i++) sp.last[i] = 0;
    for (int i = sp.dim_o; i < sp.H; i++) sp.first[i] = (uint16_t)sp.dim_l;
    for (int j = 0; j < sp.dim_i; j++) {
        const int rows = sp.target[j];
        if (j > sp.last[rows]) sp.last[rows] = (uint16_t)(j + 1);
        if (j > sp.last[rows + 1]) sp.last[rows + 1] = (uint16_t)(j + 1);
        const int rowe = sp.target[j] + sp.dim_o - 1;
        if (j < sp.first[rowe]) sp.first[rowe] = (uint16_t)j;
        if (j < sp.first[rowe + 1]) sp.first[rowe + 1] = (uint16_t)j;
    }
}

void plan_image(int dimx, int dimy, int s0, int s1, long long img, long long so_base, long long q_base, int order,
                ImagePlan& ip) {
    ip = ImagePlan();
    ip.so_total = so_base;
    ip.q_total = q_base;
    plan_append(dimx, dimy, s0, s1, img, ip, ip.range);
    ip.so_total -= so_base;
    ip.q_total -= q_base;
    plan_finish(ip, order);
}

void plan_finish(ImagePlan& ip, int order) {
    // ip.work holds rows in natural order; sort by descending cost so that a
    // dynamic scheduler hands out the long rows first.  kOrderNatural keeps
    // the images of a batch apart (image by image, each by descending L), so
    // the rows in flight read one image's oriented copies at a time.
    struct Key {
        long long img;
        double cost;
        RowWork w;
    };
    std::vector<Key> rows;
    rows.reserve(ip.work.size());
    for (const RowWork& w : ip.work) {
        const SecDesc& sd = ip.secs[w.sec];
        const uint32_t v = ip.fl[sd.rows + w.row];
        const double L = (double)((int)(v >> 16) - (int)(v & 0xffffu));
        rows.push_back({order == kOrderCostDesc ? 0 : sd.img, order == kOrderCostDesc ? L * L : L, w});
    }
    std::stable_sort(rows.begin(), rows.end(), [](const Key& x, const Key& y) {
        return x.img != y.img ? x.img < y.img : x.cost > y.cost;
    });
    for (size_t i = 0; i < rows.size(); i++) ip.work[i] = rows[i].w;
}

void plan_append(int dimx, int dimy, int s0, int s1, long long img, ImagePlan& ip, int range8[8]) {
    ip.dimx = dimx;
    ip.dimy = dimy;
    ip.s0 = s0;
    ip.s1 = s1;
    for (int t = 0; t < 8; t++) range8[t] = 0;
    int seen[4] = {0, 0, 0, 0};
    SectorPlan sp;
    long long so = ip.so_total, q = ip.q_total;
    for (int a = s0; a < s1; a++) {
        plan_sector(dimx, dimy, a, sp);
        SecDesc d{};
        d.angle = a;
        d.newAngle = sp.newAngle;
        d.type = sp.type;
        d.dim_i = sp.dim_i;
        d.dim_o = sp.dim_o;
        d.H = sp.H;
        d.offset = sp.offset;
        d.tab = (int)ip.target.size();
        d.rows = (int)ip.fl.size();
        d.stale = -1;
        // The deskew's zero-weight tap of a newAngle == 0 sector reads row H,
        // which holds what the previously processed sector left there.
        if (sp.newAngle == 0 && a > s0) d.stale = (int)ip.secs.size() - 1;
        d.pad0 = 0;
        d.R = (float)(sp.scale * sp.scale);  // (float)pow(scale,2), src/kernelSDEM.h:171
        d.rscale = (float)(1 / sp.scale);
        d.img = img;
        d.so = so;
        d.q = q;
        d.N = dimx * dimy;
        const int si = (int)ip.secs.size();
        ip.target.insert(ip.target.end(), sp.target.begin(), sp.target.begin() + sp.dim_i);
        ip.weight.insert(ip.weight.end(), sp.weight.begin(), sp.weight.begin() + sp.dim_i);
        for (int r = 0; r < sp.H; r++) {
            const int f = sp.first[r], l = sp.last[r];
            ip.fl.push_back((uint32_t)f | ((uint32_t)l << 16));
            const int L = l - f;
            if (L > 0) {
                ip.work.push_back(RowWork{si, r});
                ip.maxL = std::max(ip.maxL, L);
                ip.cells += L;
                ip.steps += (double)L * (L - 1);
            }
        }
        ip.maxH = std::max(ip.maxH, sp.H);
        so += (long long)sp.H * sp.dim_i;
        q += sp.H;
        if (!seen[sp.type]) {
            seen[sp.type] = 1;
            range8[2 * sp.type] = si;
        }
        range8[2 * sp.type + 1] = si + 1;
        ip.secs.push_back(d);
    }
    ip.so_total = so;
    ip.q_total = q;
}

void recip_table(int n, std::vector<float>& tab4) {
    tab4.assign(4 * (size_t)n + 2 * ((size_t)n + 32), 0.0f);
    float* tab2 = tab4.data() + 4 * (size_t)n;
    for (int d = 1; d < n + 32; d++) {
        const float yh = (float)(1.0 / d);
        // 1/d - yh: d*yh is exact in double and 1 - d*yh is exact (Sterbenz),
        // so the only roundings are the final double division and the float.
        const double res = (1.0 - (double)d * (double)yh) / d;
        if (d < n) {
            tab4[4 * d + 0] = yh;
            tab4[4 * d + 1] = (float)res;
            tab4[4 * d + 2] = (float)d;
            tab4[4 * d + 3] = (float)((double)d * d);
        }
        tab2[2 * d + 0] = yh;
        tab2[2 * d + 1] = (float)res;
    }
    tab4.resize((tab4.size() + 3) / 4 * 4, 0.0f);  // whole float4s
}

void fft_twiddles(int mmax, std::vector<double>& tw, std::vector<int>& off) {
    tw.clear();
    off.assign(mmax + 1, 0);
    for (int m = 1; m <= mmax; m++) {
        const unsigned N = 1u << m;
        off[m] = (int)(tw.size() / 2);
        // src/helper.h:63-81: phiT starts at e^{-i pi/N} and is squared once
        // per stage; T_l = phiT^l by repeated products.  (ac-bd, ad+bc) with
        // every product rounded, as std::complex<double> multiplies.
        double thetaT = 3.14159265358979323846264338328L / N;
        double pr = cos(thetaT), pi = -sin(thetaT);
        unsigned k = N;
        while (k > 1) {
            k >>= 1;
            double ac = pr * pr, bd = pi * pi, ad = pr * pi, bc = pi * pr;
            pr = ac - bd;
            pi = ad + bc;
            double Tr = 1.0, Ti = 0.0;
            for (unsigned l = 0; l < k; l++) {
                tw.push_back(Tr);
                tw.push_back(Ti);
                ac = Tr * pr; bd = Ti * pi; ad = Tr * pi; bc = Ti * pr;
                Tr = ac - bd;
                Ti = ad + bc;
            }
        }
    }
}

void cepstrum_table(std::vector<double>& tab) {
    static_assert(kLrtC >= kLogTab + 2 * 256 + 2, "table layout");
    tab.assign(kCtabSize, 0.0);
    for (int t = 0; t < 64; t++) {
        tab[2 * t] = cos(2.0 * M_PI * t / 64.0);
        tab[2 * t + 1] = sin(2.0 * M_PI * t / 64.0);
    }
    for (int t = 0; t < 32; t++) tab[128 + t] = cos(2.0 * M_PI * t / 32.0);
    // src/kernelRADN.h:152-162: cutoff 0.4 of the 64-sample window
    const double cutoff = 0.4, dws = 64;
    for (int k = 0; k < 26; k++) tab[160 + k] = sin((k / (cutoff * dws - 1)) * (M_PI / 2));
    // natural-log table: for the top 8 mantissa bits i, c_i = RN(1 / (1 + (i + 1/2)/256)) and
    // T_i = -log(c_i) (long double, then rounded), then ln 2 as a double-double
    for (int i = 0; i < 256; i++) {
        const double c = (double)(1.0L / (1.0L + (i + 0.5L) / 256.0L));
        tab[kLogTab + 2 * i] = c;
        tab[kLogTab + 2 * i + 1] = (double)(-logl((long double)c));
    }
    const long double ln2 = logl(2.0L);
    tab[kLogTab + 512] = (double)ln2;
    tab[kLogTab + 513] = (double)(ln2 - (long double)(double)ln2);
    // the local Radon's two DFT matrices, rows contiguous (copies of the
    // cos/sin entries above; frequencies q >= 26 stay 0)
    for (int kk = 0; kk < 64; kk++)
        for (int q = 0; q < 26; q++) tab[kLrtC + 32 * kk + q] = tab[2 * ((q * kk) & 63)];
    for (int m = 0; m < 32; m++)
        for (int q = 0; q < 26; q++) {
            const int t = (q * m) & 63;
            tab[kLrtY + 64 * m + 2 * q] = tab[2 * t];
            tab[kLrtY + 64 * m + 2 * q + 1] = tab[2 * t + 1];
        }
    // the cepstrum's 64-sample window DFT (bins k < 32), rows contiguous
    for (int m = 0; m < 64; m++)
        for (int k = 0; k < 32; k++) {
            const int t = (k * m) & 63;
            tab[kBlurD + 64 * m + 2 * k] = tab[2 * t];
            tab[kBlurD + 64 * m + 2 * k + 1] = tab[2 * t + 1];
        }
}

void lrt_gauss_table(std::vector<double>& tab) {
    // src/kernelRADN.h:123-138: sigma = winSize/5 / scale, scale = 1/cos(na)
    const double torads = M_PI / 180.0;
    tab.assign(91 * 64, 0.0);
    for (int na = 0; na <= 90; na++) {
        const double scale = 1 / cos(torads * na);
        double sigma = 64 / 5.0;
        sigma = sigma / scale;
        for (int kk = 0; kk < 64; kk++) tab[64 * na + kk] = exp(-(kk - 32) * (kk - 32) / (2 * sigma * sigma));
    }
}

}  // namespace ske
