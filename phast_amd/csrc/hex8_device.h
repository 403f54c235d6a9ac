// Device helpers shared by the voxel kernels: periodic index wrap, sum-factorised
// trilinear interpolation and gradients on the reference voxel, Gauss-point strains,
// per-voxel stiffness and closed-form 3x3 eigenvalues.
#pragma once

#include "hex8_internal.h"

namespace phast {

// Corners a = lx + 2 ly + 4 lz, Gauss points q = ix + 2 iy + 4 iz (xi fastest);
// interpolation weight of corner l at Gauss point i: al if l == i else be.

__device__ __forceinline__ int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
// periodic index for tiles wider than the cell (i in [-n, inf))
__device__ __forceinline__ int wrapm(int i, int n) { return i < 0 ? i + n : (i >= n ? i % n : i); }

// bilinear interpolation of corner values f[a][b] (a,b in {0,1}) to Gauss points o[ga][gb]
__device__ __forceinline__ void interp2(double al, double be, double f00, double f10, double f01, double f11,
                                        double o[2][2]) {
    const double t00 = al * f00 + be * f10, t10 = be * f00 + al * f10;  // [ga] at b=0
    const double t01 = al * f01 + be * f11, t11 = be * f01 + al * f11;  // [ga] at b=1
    o[0][0] = al * t00 + be * t01;
    o[0][1] = be * t00 + al * t01;
    o[1][0] = al * t10 + be * t11;
    o[1][1] = be * t10 + al * t11;
}
// transpose of interp2: corner sums R[la][lb] = sum_g W(la,ga) W(lb,gb) S[ga][gb]
__device__ __forceinline__ void interp2T(double al, double be, const double S[2][2], double R[2][2]) {
    const double t00 = al * S[0][0] + be * S[0][1], t01 = be * S[0][0] + al * S[0][1];  // [ga=0][lb]
    const double t10 = al * S[1][0] + be * S[1][1], t11 = be * S[1][0] + al * S[1][1];  // [ga=1][lb]
    R[0][0] = al * t00 + be * t10;
    R[1][0] = be * t00 + al * t10;
    R[0][1] = al * t01 + be * t11;
    R[1][1] = be * t01 + al * t11;
}

// interp2 whose second stage carries a scale: weights (as, bs) = (al, be) * h^-1
__device__ __forceinline__ void interp2s(double al, double be, double as, double bs, double f00, double f10,
                                         double f01, double f11, double o[2][2]) {
    const double t00 = al * f00 + be * f10, t10 = be * f00 + al * f10;
    const double t01 = al * f01 + be * f11, t11 = be * f01 + al * f11;
    o[0][0] = as * t00 + bs * t01;
    o[0][1] = bs * t00 + as * t01;
    o[1][0] = as * t10 + bs * t11;
    o[1][1] = bs * t10 + as * t11;
}
// transpose with scaled second stage
__device__ __forceinline__ void interp2Ts(double al, double be, double as, double bs, const double S[2][2],
                                          double R[2][2]) {
    const double t00 = al * S[0][0] + be * S[0][1], t01 = be * S[0][0] + al * S[0][1];
    const double t10 = al * S[1][0] + be * S[1][1], t11 = be * S[1][0] + al * S[1][1];
    R[0][0] = as * t00 + bs * t10;
    R[1][0] = bs * t00 + as * t10;
    R[0][1] = as * t01 + bs * t11;
    R[1][1] = bs * t01 + as * t11;
}

// Gradients of the three displacement components at the Gauss points (1/h folded into the
// second interpolation stage): gx[c][iy][iz], gy[c][ix][iz], gz[c][ix][iy]; u[c][lx+2ly+4lz].
// The first stage interpolates along the second index of the output, the second stage
// (scaled) along the first... arranged so that every derivative is the difference along its
// own axis interpolated in the two others.
__device__ __forceinline__ void voxel_grad(const Hex8Geo &G, const double u[3][8], double gx[3][2][2],
                                           double gy[3][2][2], double gz[3][2][2]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        // x-difference at (ly,lz) -> gx[iy][iz]: interp over ly (first index) scaled by 1/hx
        interp2s(G.al, G.be, G.alx, G.bex, u[c][1] - u[c][0], u[c][3] - u[c][2], u[c][5] - u[c][4],
                 u[c][7] - u[c][6], gx[c]);
        interp2s(G.al, G.be, G.aly, G.bey, u[c][2] - u[c][0], u[c][3] - u[c][1], u[c][6] - u[c][4],
                 u[c][7] - u[c][5], gy[c]);
        interp2s(G.al, G.be, G.alz, G.bez, u[c][4] - u[c][0], u[c][5] - u[c][1], u[c][6] - u[c][2],
                 u[c][7] - u[c][3], gz[c]);
    }
}

// engineering strain (xx,yy,zz,yz,xz,xy) at Gauss point (ix,iy,iz); the macro strain has
// been added into the gradient sets beforehand (fold_macro).
__device__ __forceinline__ void qp_strain(const double gx[3][2][2], const double gy[3][2][2],
                                          const double gz[3][2][2], int ix, int iy, int iz, double e[6]) {
    e[0] = gx[0][iy][iz];
    e[1] = gy[1][ix][iz];
    e[2] = gz[2][ix][iy];
    e[3] = gz[1][ix][iy] + gy[2][ix][iz];
    e[4] = gz[0][ix][iy] + gx[2][iy][iz];
    e[5] = gy[0][ix][iz] + gx[1][iy][iz];
}

// add E into the gradient entries each strain component starts from
__device__ __forceinline__ void fold_macro(const double E[6], double gx[3][2][2], double gy[3][2][2],
                                           double gz[3][2][2]) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            gx[0][a][b] += E[0];
            gy[1][a][b] += E[1];
            gz[2][a][b] += E[2];
            gz[1][a][b] += E[3];
            gz[0][a][b] += E[4];
            gy[0][a][b] += E[5];
        }
}

// trilinear interpolation of corner scalars to the 8 Gauss points (xi fastest)
__device__ __forceinline__ void voxel_interp(const Hex8Geo &G, const double d[8], double dq[8]) {
    double ax[2][2][2];  // [ly][lz][ix]
#pragma unroll
    for (int ly = 0; ly < 2; ++ly)
#pragma unroll
        for (int lz = 0; lz < 2; ++lz) {
            const double c0 = d[2 * ly + 4 * lz], c1 = d[1 + 2 * ly + 4 * lz];
            ax[ly][lz][0] = G.al * c0 + G.be * c1;
            ax[ly][lz][1] = G.be * c0 + G.al * c1;
        }
#pragma unroll
    for (int ix = 0; ix < 2; ++ix) {
        double o[2][2];
        interp2(G.al, G.be, ax[0][0][ix], ax[1][0][ix], ax[0][1][ix], ax[1][1][ix], o);
#pragma unroll
        for (int iy = 0; iy < 2; ++iy)
#pragma unroll
            for (int iz = 0; iz < 2; ++iz) dq[ix + 2 * iy + 4 * iz] = o[iy][iz];
    }
}

// Engineering-Voigt stiffness of voxel e: from (lam, mu) when every bank entry is isotropic.
__device__ __forceinline__ void voxel_C(bool iso, const double *bank, const int32_t *bank_index, const double *lam,
                                        const double *mu, int64_t e, double C[36]) {
    if (iso) {
        const double l = lam[e], m = mu[e];
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c)
                C[r * 6 + c] = ((r < 3 && c < 3) ? l : 0.0) + (r == c ? (r < 3 ? 2.0 * m : m) : 0.0);
    } else {
        const double *B = bank + 36 * (int64_t)bank_index[e];
#pragma unroll
        for (int k = 0; k < 36; ++k) C[k] = B[k];
    }
}

// Closed-form ascending eigenvalues of a symmetric 3x3 (splits.py:93-125).
__device__ __forceinline__ void eig3(double a11, double a22, double a33, double a12, double a13, double a23,
                                     double ev[3]) {
    const double q = (a11 + a22 + a33) / 3.0;
    const double b11 = a11 - q, b22 = a22 - q, b33 = a33 - q;
    const double p1 = a12 * a12 + a13 * a13 + a23 * a23;
    const double p = sqrt((b11 * b11 + b22 * b22 + b33 * b33 + 2.0 * p1) / 6.0);
    const double det_b = b11 * (b22 * b33 - a23 * a23) - a12 * (a12 * b33 - a23 * a13) + a13 * (a12 * a23 - b22 * a13);
    const bool iso = p <= 2.2250738585072014e-308;
    const double ps = iso ? 1.0 : p;
    double r = det_b / (2.0 * (ps * ps * ps));
    r = fmin(fmax(r, -1.0), 1.0);
    const double phi = acos(r) / 3.0;
    const double largest = q + 2.0 * p * cos(phi);
    const double smallest = q + 2.0 * p * cos(phi + 2.0 * M_PI / 3.0);
    const double middle = 3.0 * q - largest - smallest;
    if (iso) {
        ev[0] = ev[1] = ev[2] = q;
    } else {
        ev[0] = smallest;
        ev[1] = middle;
        ev[2] = largest;
    }
}

// Internal force of one voxel: r[c][corner] = sum_q w B_q^T (g_q C (B_q u + E))
// (operator.py:1108-1169).  With OWN, sE += sum_q w sigma_q and (ENERGY) en += 0.5 sum_q
// w sigma.eps.  sum-factorised gradients; g w folded into the moduli.
//
// Crack closure (CL != 0; operator.py:1039-1105, splits.py:424-742):
// sigma += (1 - g) sum_k s_k k_k (v_k . eps) v_k.  Isotropic cells (ISO): the rows
// v = (1, e_x, e_y, e_z) with moduli (lamc, 2 muc, 2 muc, 2 muc) -- zero where the voxel adds
// nothing back.  General banks: the four rows and moduli of the voxel's bank entry in
// T[0..23], T[24..27] (the cell axes of an isotropic phase or the frame of a laminate band),
// where `clon`.  The bits s_k are the signs [v_k . eps < 0] of the voxel's own strain
// (CL == 1) or bit 4 q + k of `pat` (CL == 2).
template <bool ISO, bool ENERGY, int CL = 0>
__device__ __forceinline__ void voxel_mech(const Hex8Geo &G, const double u[3][8], const double E[6],
                                           const double g[8], const double *__restrict__ C, double lam,
                                           double mu, bool own, double r[3][8], double sE[6], double &en,
                                           double lamc = 0.0, double muc = 0.0, uint32_t pat = 0u,
                                           const double *__restrict__ T = nullptr, bool clon = false) {
    double gx[3][2][2], gy[3][2][2], gz[3][2][2];
    voxel_grad(G, u, gx, gy, gz);
    fold_macro(E, gx, gy, gz);
    // each partial sum takes two Gauss points: the first assigns, the second adds
    double Sx[3][2][2], Sy[3][2][2], Sz[3][2][2];
    double ew = 0.0;
#pragma unroll
    for (int iz = 0; iz < 2; ++iz)
#pragma unroll
        for (int iy = 0; iy < 2; ++iy)
#pragma unroll
            for (int ix = 0; ix < 2; ++ix) {
                double e[6];
                qp_strain(gx, gy, gz, ix, iy, iz, e);
                const double gw = g[ix + 2 * iy + 4 * iz] * G.w;
                double s[6];
                if (ISO) {
                    const double lg = lam * gw, mg = mu * gw, mg2 = 2.0 * mg;
                    const double lt = lg * ((e[0] + e[1]) + e[2]);
                    s[0] = lt + mg2 * e[0];
                    s[1] = lt + mg2 * e[1];
                    s[2] = lt + mg2 * e[2];
                    s[3] = mg * e[3];
                    s[4] = mg * e[4];
                    s[5] = mg * e[5];
                    if (CL != 0) {
                        const int q = ix + 2 * iy + 4 * iz;
                        const double cw = (1.0 - g[q]) * G.w;
                        const double tr = (e[0] + e[1]) + e[2];
                        bool b0, b1, b2, b3;
                        if (CL == 1) {
                            b0 = tr < 0.0;
                            b1 = e[0] < 0.0;
                            b2 = e[1] < 0.0;
                            b3 = e[2] < 0.0;
                        } else {
                            const uint32_t bits = pat >> (4 * q);
                            b0 = bits & 1u;
                            b1 = bits & 2u;
                            b2 = bits & 4u;
                            b3 = bits & 8u;
                        }
                        const double ct = b0 ? lamc * tr : 0.0, m2 = 2.0 * muc;
                        s[0] += cw * (ct + (b1 ? m2 * e[0] : 0.0));
                        s[1] += cw * (ct + (b2 ? m2 * e[1] : 0.0));
                        s[2] += cw * (ct + (b3 ? m2 * e[2] : 0.0));
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        double t = 0.0;
#pragma unroll
                        for (int m = 0; m < 6; ++m) t += C[k * 6 + m] * e[m];
                        s[k] = gw * t;
                    }
                    if (CL != 0 && clon) {
                        const int q = ix + 2 * iy + 4 * iz;
                        const double cw = (1.0 - g[q]) * G.w;
                        double c4[4];
#pragma unroll
                        for (int b = 0; b < 4; ++b) {
                            double v = 0.0;
#pragma unroll
                            for (int m = 0; m < 6; ++m) v += T[6 * b + m] * e[m];
                            const bool act = CL == 1 ? v < 0.0 : ((pat >> (4 * q + b)) & 1u) != 0u;
                            c4[b] = act ? T[24 + b] * v : 0.0;
                        }
#pragma unroll
                        for (int m = 0; m < 6; ++m)
                            s[m] += cw * (((c4[0] * T[m] + c4[1] * T[6 + m]) + c4[2] * T[12 + m]) + c4[3] * T[18 + m]);
                    }
                }
                if (ENERGY) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < 6; ++k) t += s[k] * e[k];
                    ew += t;
                }
                if (ix == 0) {
                    Sx[0][iy][iz] = s[0];
                    Sx[1][iy][iz] = s[5];
                    Sx[2][iy][iz] = s[4];
                } else {
                    Sx[0][iy][iz] += s[0];
                    Sx[1][iy][iz] += s[5];
                    Sx[2][iy][iz] += s[4];
                }
                if (iy == 0) {
                    Sy[0][ix][iz] = s[5];
                    Sy[1][ix][iz] = s[1];
                    Sy[2][ix][iz] = s[3];
                } else {
                    Sy[0][ix][iz] += s[5];
                    Sy[1][ix][iz] += s[1];
                    Sy[2][ix][iz] += s[3];
                }
                if (iz == 0) {
                    Sz[0][ix][iy] = s[4];
                    Sz[1][ix][iy] = s[3];
                    Sz[2][ix][iy] = s[2];
                } else {
                    Sz[0][ix][iy] += s[4];
                    Sz[1][ix][iy] += s[3];
                    Sz[2][ix][iy] += s[2];
                }
            }
    if (own) {  // sum_q w sigma_q from the partial sums (each component appears once per channel)
        sE[0] += (Sx[0][0][0] + Sx[0][1][0]) + (Sx[0][0][1] + Sx[0][1][1]);
        sE[1] += (Sy[1][0][0] + Sy[1][1][0]) + (Sy[1][0][1] + Sy[1][1][1]);
        sE[2] += (Sz[2][0][0] + Sz[2][1][0]) + (Sz[2][0][1] + Sz[2][1][1]);
        sE[3] += (Sy[2][0][0] + Sy[2][1][0]) + (Sy[2][0][1] + Sy[2][1][1]);
        sE[4] += (Sx[2][0][0] + Sx[2][1][0]) + (Sx[2][0][1] + Sx[2][1][1]);
        sE[5] += (Sx[1][0][0] + Sx[1][1][0]) + (Sx[1][0][1] + Sx[1][1][1]);
        if (ENERGY) en += 0.5 * ew;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double R[2][2];
        interp2Ts(G.al, G.be, G.alx, G.bex, Sx[c], R);  // R[ly][lz]
#pragma unroll
        for (int ly = 0; ly < 2; ++ly)
#pragma unroll
            for (int lz = 0; lz < 2; ++lz) {
                r[c][1 + 2 * ly + 4 * lz] = R[ly][lz];
                r[c][2 * ly + 4 * lz] = -R[ly][lz];
            }
        interp2Ts(G.al, G.be, G.aly, G.bey, Sy[c], R);  // R[lx][lz]
#pragma unroll
        for (int lx = 0; lx < 2; ++lx)
#pragma unroll
            for (int lz = 0; lz < 2; ++lz) {
                r[c][lx + 2 + 4 * lz] += R[lx][lz];
                r[c][lx + 4 * lz] -= R[lx][lz];
            }
        interp2Ts(G.al, G.be, G.alz, G.bez, Sz[c], R);  // R[lx][ly]
#pragma unroll
        for (int lx = 0; lx < 2; ++lx)
#pragma unroll
            for (int ly = 0; ly < 2; ++ly) {
                r[c][lx + 2 * ly + 4] += R[lx][ly];
                r[c][lx + 2 * ly] -= R[lx][ly];
            }
    }
}

}  // namespace phast
