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

// Walsh-Hadamard modes of the trilinear field with corner values u: W_S = sum_a s_S(a) u[a]
// with s_x(a) = 2 lx - 1 etc., so that u = (1/8) sum_S W_S x^Sx y^Sy z^Sz on the centred
// voxel [-1,1]^3.  W = {x, y, z, xy, xz, yz, xyz}; the constant mode is not formed.
__device__ __forceinline__ void voxel_modes(const double u[8], double W[7]) {
    double s[4], d[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {  // j = ly + 2 lz
        s[j] = u[2 * j + 1] + u[2 * j];
        d[j] = u[2 * j + 1] - u[2 * j];
    }
    const double ss0 = s[1] + s[0], ss1 = s[3] + s[2], sd0 = s[1] - s[0], sd1 = s[3] - s[2];
    const double ds0 = d[1] + d[0], ds1 = d[3] + d[2], dd0 = d[1] - d[0], dd1 = d[3] - d[2];
    W[0] = ds1 + ds0;
    W[1] = sd1 + sd0;
    W[2] = ss1 - ss0;
    W[3] = dd1 + dd0;
    W[4] = ds1 - ds0;
    W[5] = sd1 - sd0;
    W[6] = dd1 - dd0;
}

// voxel_grad followed by fold_macro, from the modes: at the Gauss point (sx, sy, sz) c3
// du/dX = kx (W_x + c3 sy W_xy + c3 sz W_xz + c3^2 sy sz W_xyz), kx = 1/(4 hx), and alike for
// y and z; the y and z derivatives share W_yz + c3 sx W_xyz.  The macro strain entries are
// added to the constant terms kx W_x, ky W_y, kz W_z once instead of at four points each.
__device__ __forceinline__ void voxel_grad_macro(const Hex8Geo &G, const double u[3][8], const double E[6],
                                                 double gx[3][2][2], double gy[3][2][2], double gz[3][2][2]) {
    const double c = G.c3;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double W[7];
        voxel_modes(u[k], W);
        double bx = G.kx * W[0], by = G.ky * W[1], bz = G.kz * W[2];
        if (k == 0) {  // gx[0] += E0, gy[0] += E5, gz[0] += E4
            bx += E[0];
            by += E[5];
            bz += E[4];
        } else if (k == 1) {  // gy[1] += E1, gz[1] += E3
            by += E[1];
            bz += E[3];
        } else {  // gz[2] += E2
            bz += E[2];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const double si = i ? 1.0 : -1.0;
            const double X = fma(si * G.kcx, W[3], bx);  // x derivative at sy = si
            const double Y = fma(si * c, W[6], W[4]);
            gx[k][i][0] = fma(-G.kcx, Y, X);
            gx[k][i][1] = fma(G.kcx, Y, X);
            const double V = fma(si * c, W[6], W[5]);  // shared by y and z at sx = si
            const double Xy = fma(si * G.kcy, W[3], by);
            gy[k][i][0] = fma(-G.kcy, V, Xy);
            gy[k][i][1] = fma(G.kcy, V, Xy);
            const double Xz = fma(si * G.kcz, W[4], bz);
            gz[k][i][0] = fma(-G.kcz, V, Xz);
            gz[k][i][1] = fma(G.kcz, V, Xz);
        }
    }
}

// Transpose of voxel_grad_macro without the macro terms: corner forces
// r[a] = sum_q (dN_a/dX Sx + dN_a/dY Sy + dN_a/dZ Sz)(q) for the Gauss-point sums
// Sx[iy][iz] (already summed over ix), Sy[ix][iz], Sz[ix][iy] of one component.  The sums
// T of Sx, Sy, Sz over their four points (the macro stress rows) are returned in Tsum.
__device__ __forceinline__ void voxel_gradT(const Hex8Geo &G, const double Sx[2][2], const double Sy[2][2],
                                            const double Sz[2][2], double r[8], double Tsum[3]) {
    const double c = G.c3;
    // 2x2 butterflies: (sum, second-index sign, first-index sign, both) of each direction
    const double ax0 = Sx[1][0] + Sx[0][0], ax1 = Sx[1][1] + Sx[0][1];
    const double bx0 = Sx[1][0] - Sx[0][0], bx1 = Sx[1][1] - Sx[0][1];
    const double ay0 = Sy[1][0] + Sy[0][0], ay1 = Sy[1][1] + Sy[0][1];
    const double by0 = Sy[1][0] - Sy[0][0], by1 = Sy[1][1] - Sy[0][1];
    const double az0 = Sz[1][0] + Sz[0][0], az1 = Sz[1][1] + Sz[0][1];
    const double bz0 = Sz[1][0] - Sz[0][0], bz1 = Sz[1][1] - Sz[0][1];
    Tsum[0] = ax1 + ax0;
    Tsum[1] = ay1 + ay0;
    Tsum[2] = az1 + az0;
    // force modes: F_x = kx T_x ...; the mixed ones before their c3 factors
    const double Fx = G.kx * Tsum[0], Fy = G.ky * Tsum[1], Fz = G.kz * Tsum[2];
    const double Gxy = fma(G.kx, bx1 + bx0, G.ky * (by1 + by0));
    const double Gxz = fma(G.kx, ax1 - ax0, G.kz * (bz1 + bz0));
    const double Gyz = fma(G.ky, ay1 - ay0, G.kz * (az1 - az0));
    const double Gxyz = fma(G.kx, bx1 - bx0, fma(G.ky, by1 - by0, G.kz * (bz1 - bz0)));
    // r(sx,sy,sz) = sx Fx + sy Fy + sz Fz + c3 (sx sy Gxy + sx sz Gxz + sy sz Gyz) + c3^2 sx sy sz Gxyz
#pragma unroll
    for (int ly = 0; ly < 2; ++ly) {
        const double sy = ly ? 1.0 : -1.0;
        const double U = fma(sy * c, Gxy, Fx);
        const double Z = fma(sy * c, Gyz, Fz);
        const double V = fma(sy * c, Gxyz, Gxz);
        const double P0 = sy * Fy - U, P1 = sy * Fy + U;  // lx = 0, 1
        const double Q0 = fma(-c, V, Z), Q1 = fma(c, V, Z);
        r[2 * ly] = P0 - Q0;
        r[2 * ly + 1] = P1 - Q1;
        r[2 * ly + 4] = P0 + Q0;
        r[2 * ly + 5] = P1 + Q1;
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

// The 8 Gauss-point values of voxel e from an (E, 8) array: four 16-byte loads when the
// array is 16-byte aligned (a wave's loads then touch each 64-byte row once per pair of
// values instead of once per value), else eight 8-byte loads.  The branch is uniform.
__device__ __forceinline__ void load_qp8(const double *__restrict__ a, int64_t e, double v[8]) {
    if ((reinterpret_cast<uintptr_t>(a) & 15) == 0) {
        const double2 *p = reinterpret_cast<const double2 *>(a + e * 8);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double2 t = p[k];
            v[2 * k] = t.x;
            v[2 * k + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = a[e * 8 + q];
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
    voxel_grad_macro(G, u, E, gx, gy, gz);
    // each partial sum takes two Gauss points: the first assigns, the second adds
    double Sx[3][2][2], Sy[3][2][2], Sz[3][2][2];
    double ew = 0.0;
    const double lw = lam * G.w, mw = mu * G.w, mw2 = 2.0 * mw;  // (ISO) moduli times the weight
#pragma unroll
    for (int iz = 0; iz < 2; ++iz)
#pragma unroll
        for (int iy = 0; iy < 2; ++iy)
#pragma unroll
            for (int ix = 0; ix < 2; ++ix) {
                double e[6];
                qp_strain(gx, gy, gz, ix, iy, iz, e);
                double s[6];
                if (ISO) {
                    const double gq = g[ix + 2 * iy + 4 * iz];
                    const double lg = lw * gq, mg = mw * gq, mg2 = mw2 * gq;
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
                    const double gw = g[ix + 2 * iy + 4 * iz] * G.w;
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
    double Ts[3][3];  // Ts[k][dir]: sum over the four points of the direction's sums
#pragma unroll
    for (int k = 0; k < 3; ++k) voxel_gradT(G, Sx[k], Sy[k], Sz[k], r[k], Ts[k]);
    if (own) {  // sum_q w sigma_q: each component appears once among the sums
        sE[0] += Ts[0][0];
        sE[1] += Ts[1][1];
        sE[2] += Ts[2][2];
        sE[3] += Ts[2][1];
        sE[4] += Ts[2][0];
        sE[5] += Ts[1][0];
        if (ENERGY) en += 0.5 * ew;
    }
}

}  // namespace phast
