// Constant-source panel potential of the ibctyp = 3 boundary sums.
//
//   PanelFrame : the frame calc_boundinf_3d_s / calc_boundinf_3d_lev_s set up for the TARGET
//                plane (pmbound.f90:537-620, 1271-1337).  Every source panel is integrated
//                as if it lay in that plane, with its extents S (along E1) and T (along E2);
//                on X planes the extents are (DY, DZ) while E1 runs along z — the panel is
//                turned a quarter when DY != DZ, as in the reference.
//   fsour_a4   : FSOUR_A4 (pmbound.f90:1541-1632) with NSIDE = 0, ISING = 0, EPSS = 1e-14.
//
// The local coordinates X, Y, Z are signed copies of the global offsets (the frame vectors
// are signed axes), taken in the reference's order so RO, the far-field test RO/DIAG > 10
// and every branch decision of the near field come out as in the reference.
#pragma once
#include <hip/hip_runtime.h>

namespace vpmx {

struct PanelFrame {
    int a1, a2, an;        // global axes of E1, E2, E3
    double s1, s2, sn;     // their signs
    double S[4], T[4], D[4];
    double DIAG, AREA;
};

// iplane: +-1 Y planes, +-2 X planes, +-3 Z planes; negative on the max faces.
__host__ __device__ inline PanelFrame panel_frame(int iplane, const double h[3]) {
    PanelFrame f;
    const double si = iplane > 0 ? 1.0 : -1.0;
    const double DX = h[0], DY = h[1], DZ = h[2];
    const int ap = iplane < 0 ? -iplane : iplane;
    double a, b;
    if (ap == 1) {  // E1 = si x, E2 = z, E3 = -si y
        f.DIAG = sqrt(DX * DX + DZ * DZ);
        f.a1 = 0; f.s1 = 1.0 * si; f.a2 = 2; f.s2 = 1.0; f.an = 1; f.sn = -1.0 * si;
        f.AREA = DX * DZ; a = DX; b = DZ;
    } else if (ap == 2) {  // E1 = z, E2 = -si y, E3 = -si x
        f.DIAG = sqrt(DY * DY + DZ * DZ);
        f.a1 = 2; f.s1 = 1.0; f.a2 = 1; f.s2 = -1.0 * si; f.an = 0; f.sn = -1.0 * si;
        f.AREA = DY * DZ; a = DY; b = DZ;
    } else {  // E1 = x, E2 = -si y, E3 = -si z
        f.DIAG = sqrt(DX * DX + DY * DY);
        f.a1 = 0; f.s1 = 1.0; f.a2 = 1; f.s2 = -1.0 * si; f.an = 2; f.sn = -1.0 * si;
        f.AREA = DX * DY; a = DX; b = DY;
    }
    f.S[0] = -0.5 * a; f.S[1] = -0.5 * a; f.S[2] = 0.5 * a; f.S[3] = 0.5 * a;
    f.T[0] = -0.5 * b; f.T[1] = 0.5 * b; f.T[2] = 0.5 * b; f.T[3] = -0.5 * b;
    f.D[0] = b; f.D[1] = a; f.D[2] = b; f.D[3] = a;
    return f;
}

__host__ __device__ inline double pick3(const double v[3], int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }

// Local coordinates of XO - RG in the frame, and RO.
__host__ __device__ inline void panel_local(const PanelFrame &f, const double d[3], double *X, double *Y, double *Z,
                                            double *RO) {
#pragma clang fp contract(off)
    *X = f.s1 * pick3(d, f.a1);
    *Y = f.s2 * pick3(d, f.a2);
    *Z = f.sn * pick3(d, f.an);
    *RO = sqrt(*X * *X + *Y * *Y + *Z * *Z);
}

// Near-field value of FSOUR_A4 at local (X, Y, Z) (the caller has tested RO / DIAG <= 10).
// SINB = (1, 0, -1, 0), COSB = (0, 1, 0, -1): the angle terms exist for edges 2 and 4 only.
__host__ __device__ inline double fsour_near(const PanelFrame &f, double X, double Y, double Z, double PI4) {
#pragma clang fp contract(off)
    const double EPSS = 1e-14, TINY = 1e-10;
    const double SINB[4] = {1.0, 0.0, -1.0, 0.0}, COSB[4] = {0.0, 1.0, 0.0, -1.0};
    const double ZP = Z < 0.0 ? -1.0 : 1.0;
    double AZ = fabs(Z);
    AZ = AZ > EPSS ? AZ : EPSS;
    Z = AZ * ZP;
    const double ZK = Z * Z;
    double R[4], E[4], H[4];
#pragma unroll
    for (int K = 0; K < 4; K++) {
        const double XK = (X - f.S[K]) * (X - f.S[K]);
        const double YK = (Y - f.T[K]) * (Y - f.T[K]);
        H[K] = (X - f.S[K]) * (Y - f.T[K]);
        R[K] = sqrt(XK + YK + ZK);
        E[K] = ZK + XK;
    }
    double FIL = 0.0;
#pragma unroll
    for (int K = 0; K < 4; K++) {
        const int K2 = K == 3 ? 0 : K + 1;
        const double A2 = R[K] + R[K2] + f.D[K];
        const double A1 = R[K] + R[K2] - f.D[K];
        if (fabs(A1 * A2) < TINY) continue;
        const double AK = log(A1 / A2);
        const double TOO = -(X - f.S[K]) * SINB[K] + (Y - f.T[K]) * COSB[K];
        FIL = FIL - TOO * AK;
    }
    AZ = AZ > TINY ? AZ : TINY;
    Z = AZ * ZP;
    double UL3 = 0.0;
#pragma unroll
    for (int K = 1; K < 4; K += 2) {
        const int K2 = K == 3 ? 0 : K + 1;
        UL3 = UL3 + atan((SINB[K] * E[K] - COSB[K] * H[K]) / (Z * COSB[K] * R[K])) -
              atan((SINB[K] * E[K2] - COSB[K] * H[K2]) / (Z * COSB[K] * R[K2]));
    }
    return 0.0 + (FIL + Z * UL3) / PI4;
}

// FSOUR_A4 with panel area `area` (the frame's AREA in the direct sums, a cell group's
// area in the tree).
__host__ __device__ inline double fsour_a4(const PanelFrame &f, const double d[3], double area, double PI4) {
#pragma clang fp contract(off)
    double X, Y, Z, RO;
    panel_local(f, d, &X, &Y, &Z, &RO);
    const double RATIO = RO / f.DIAG;
    if (RATIO > 10) {
        const double AREA1 = area / PI4;
        return -AREA1 / RO;
    }
    return fsour_near(f, X, Y, Z, PI4);
}

}  // namespace vpmx
