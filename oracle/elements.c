/*
 * ORACLE — test infrastructure only.  Nothing in the product library links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 *
 * Plain-C restatement of the FEDEM Femlib element routines on the Part
 * Reducer hot path.  Every function cites the reference routine it follows.
 * Matrices are Fortran column-major: A[i + j*n].
 *
 * Parity pins (see tests/test_oracle_elements.py):
 *   hex8_stiff   -> oracle/_ref build of Femlib/hexa.f HEXA31 (bitwise-close)
 *   hex8_mass    -> closed form of the trilinear brick mass, total mass,
 *                   row-sum lumping; HEXA35 itself needs BLAS DGEMM that the
 *                   image lacks, so no _ref pin (closed form + invariants)
 */
#include <math.h>
#include <string.h>

/* Femlib/isoMatModule.f90:63-85 (isoMat3D) */
void iso_mat3d(double E, double nu, double C[36])
{
    memset(C, 0, 36 * sizeof(double));
    double fac = E / ((1.0 + nu) * (1.0 - nu - nu));
    C[0] = (1.0 - nu) * fac;
    C[1] = nu * fac; C[2] = C[1];
    C[6] = C[1]; C[7] = C[0]; C[8] = C[1];
    C[12] = C[1]; C[13] = C[1]; C[14] = C[0];
    C[21] = (0.5 - nu) * fac; C[28] = C[21]; C[35] = C[21];
}

/* Femlib/isoMatModule.f90:21-32 (isoMat2D), 3x3 */
void iso_mat2d(double E, double nu, double C[9])
{
    memset(C, 0, 9 * sizeof(double));
    C[0] = E / (1.0 - nu * nu);
    C[3] = nu * C[0]; C[1] = C[3]; C[4] = C[0];
    C[8] = 0.5 * E / (1.0 + nu);
}

/* Femlib/hexa.f:802-870 (JABN30): Jacobian at (xi,eta,zeta) and its inverse.
 * Returns detJ; BJ is the inverse Jacobian (only when want_inv). */
static double jabn30(double BJ[9], const double *X, const double *Y, const double *Z,
                     double xi, double eta, double zeta, int want_inv)
{
    double A1 = 1. + xi, A2 = 1. - xi, B1 = 1. + eta, B2 = 1. - eta, C1 = 1. + zeta, C2 = 1. - zeta;
#define JROW1(V) (0.125 * (B1 * C1 * (V[6] - V[7]) - B2 * C1 * (V[4] - V[5]) + B1 * C2 * (V[2] - V[3]) - B2 * C2 * (V[0] - V[1])))
#define JROW2(V) (0.125 * (A1 * C1 * (V[6] - V[5]) + A2 * C1 * (V[7] - V[4]) + A1 * C2 * (V[2] - V[1]) + A2 * C2 * (V[3] - V[0])))
#define JROW3(V) (0.125 * (A1 * B1 * (V[6] - V[2]) + A2 * B1 * (V[7] - V[3]) + A2 * B2 * (V[4] - V[0]) + A1 * B2 * (V[5] - V[1])))
    double J11 = JROW1(X), J12 = JROW1(Y), J13 = JROW1(Z);
    double J21 = JROW2(X), J22 = JROW2(Y), J23 = JROW2(Z);
    double J31 = JROW3(X), J32 = JROW3(Y), J33 = JROW3(Z);
#undef JROW1
#undef JROW2
#undef JROW3
    double detJ = J11 * (J22 * J33 - J23 * J32) - J12 * (J21 * J33 - J23 * J31) + J13 * (J21 * J32 - J22 * J31);
    if (want_inv && detJ > 0.0) {
        double det = 1. / detJ;
        /* BJ(i,j) column-major, BJ(I,J) at [I-1 + 3*(J-1)] */
        BJ[0] = det * (J22 * J33 - J32 * J23);
        BJ[3] = -det * (J12 * J33 - J13 * J32);
        BJ[6] = det * (J12 * J23 - J13 * J22);
        BJ[1] = -det * (J21 * J33 - J23 * J31);
        BJ[4] = det * (J11 * J33 - J13 * J31);
        BJ[7] = -det * (J11 * J23 - J13 * J21);
        BJ[2] = det * (J21 * J32 - J22 * J31);
        BJ[5] = -det * (J11 * J32 - J12 * J31);
        BJ[8] = det * (J11 * J22 - J12 * J21);
    }
    return detJ;
}

/* Femlib/hexa.f:7-270 (HEXA31), NIP=3, IOP=0 (no incompatible modes), which is
 * what ADD44 uses by default (inadd.f90:2148-2155).  The Gauss abscissa and
 * weights are default-REAL literals in the Fortran source (SQRT(0.6), 5.0/9.0,
 * 8.0/9.0), i.e. single precision rounded, then widened; reproduced here.
 * Returns 0, or -2 for a non-positive nodal Jacobian. */
int hex8_stiff(const double X[8], const double Y[8], const double Z[8],
               double E, double nu, double EK[576])
{
    static const double cxi[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const double ceta[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const double czeta[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double C[36], BJ[9], D[3 * 8], EP[6 * 24], EKT[576];
    iso_mat3d(E, nu, C);
    for (int i = 0; i < 8; i++)
        if (jabn30(BJ, X, Y, Z, cxi[i], ceta[i], czeta[i], 0) < 0.0) return -2;
    double abc[3], w[3];
    abc[0] = -(double)sqrtf(0.6f); abc[1] = 0.0; abc[2] = -abc[0];
    w[0] = (double)(5.0f / 9.0f); w[1] = (double)(8.0f / 9.0f); w[2] = w[0];
    memset(EKT, 0, sizeof EKT);
#define Cm(i, j) C[(i) + 6 * (j)]
#define Dm(n, m) D[(n) + 3 * (m)]
#define EPm(n, c) EP[(n) + 6 * (c)]
#define BJm(i, j) BJ[(i) + 3 * (j)]
    for (int k = 0; k < 3; k++) {
        double zeta = abc[k], zetp = 1. + zeta, zetm = 1. - zeta, wz = w[k];
        for (int j = 0; j < 3; j++) {
            double eta = abc[j], etap = 1. + eta, etam = 1. - eta, wy = w[j];
            for (int i = 0; i < 3; i++) {
                double xi = abc[i], xip = 1. + xi, xim = 1. - xi;
                double wi = wz * wy * w[i];
                double detJ = jabn30(BJ, X, Y, Z, xi, eta, zeta, 1);
                for (int n = 0; n < 3; n++) {
                    double b1 = BJm(n, 0), b2 = BJm(n, 1), b3 = BJm(n, 2);
                    Dm(n, 0) = .125 * (-b1 * etam * zetm - b2 * xim * zetm - b3 * xim * etam);
                    Dm(n, 1) = .125 * (b1 * etam * zetm - b2 * xip * zetm - b3 * xip * etam);
                    Dm(n, 2) = .125 * (b1 * etap * zetm + b2 * xip * zetm - b3 * xip * etap);
                    Dm(n, 3) = .125 * (-b1 * etap * zetm + b2 * xim * zetm - b3 * xim * etap);
                    Dm(n, 4) = .125 * (-b1 * etam * zetp - b2 * xim * zetp + b3 * xim * etam);
                    Dm(n, 5) = .125 * (b1 * etam * zetp - b2 * xip * zetp + b3 * xip * etam);
                    Dm(n, 6) = .125 * (b1 * etap * zetp + b2 * xip * zetp + b3 * xip * etap);
                    Dm(n, 7) = .125 * (-b1 * etap * zetp + b2 * xim * zetp + b3 * xim * etap);
                }
                for (int m = 0; m < 8; m++)
                    for (int n = 0; n < 6; n++) {
                        EPm(n, 3 * m) = detJ * (Cm(n, 0) * Dm(0, m) + Cm(n, 3) * Dm(1, m) + Cm(n, 4) * Dm(2, m));
                        EPm(n, 3 * m + 1) = detJ * (Cm(n, 1) * Dm(1, m) + Cm(n, 3) * Dm(0, m) + Cm(n, 5) * Dm(2, m));
                        EPm(n, 3 * m + 2) = detJ * (Cm(n, 2) * Dm(2, m) + Cm(n, 4) * Dm(0, m) + Cm(n, 5) * Dm(1, m));
                    }
                for (int l = 0; l < 8; l++) {
                    int m = 3 * l;
                    for (int n = m; n < 24; n++) {
                        EKT[m + 24 * n] += wi * (Dm(0, l) * EPm(0, n) + Dm(1, l) * EPm(3, n) + Dm(2, l) * EPm(4, n));
                        EKT[m + 1 + 24 * n] += wi * (Dm(1, l) * EPm(1, n) + Dm(0, l) * EPm(3, n) + Dm(2, l) * EPm(5, n));
                        EKT[m + 2 + 24 * n] += wi * (Dm(2, l) * EPm(2, n) + Dm(0, l) * EPm(4, n) + Dm(1, l) * EPm(5, n));
                    }
                }
            }
        }
    }
#undef Cm
#undef Dm
#undef EPm
#undef BJm
    /* lower part from symmetry (hexa.f:184-186): the loop writes the upper
     * triangle including the within-block lower entries of diagonal blocks,
     * then EKT(J,I)=EKT(I,J) for J>=I overwrites the lower triangle. */
    for (int i = 0; i < 24; i++)
        for (int j = i; j < 24; j++) EKT[j + 24 * i] = EKT[i + 24 * j];
    memcpy(EK, EKT, sizeof EKT);
    return 0;
}

/* Femlib/hexa.f:1116-1140 (LINHEX) shape functions / derivatives. */
static void linhex(int g, double xsi, double eta, double zeta, double *W, int ider)
{
    static const double X[8] = {1., -1., -1., 1., 1., -1., -1., 1.};
    static const double Y[8] = {1., 1., -1., -1., 1., 1., -1., -1.};
    static const double Z[8] = {1., 1., 1., 1., -1., -1., -1., -1.};
    double A = 1.0 + xsi * X[g], B = 1.0 + eta * Y[g], C = 1.0 + zeta * Z[g];
    if (ider <= 0) {
        W[0] = 0.125 * A * B * C;
    } else {
        W[0] = 0.125 * X[g] * B * C;
        W[1] = 0.125 * Y[g] * A * C;
        W[2] = 0.125 * Z[g] * A * B;
    }
}

/* Femlib/hexa.f:1162-1224 (GAQDTA) for NPOINT=3: the DATA tables are default
 * REAL literals, so abscissa and weight are single-precision rounded. */
static void gaqdta3(int i, double *a, double *w)
{
    const double a3 = (double)0.77459666924148f;
    const double w1 = (double)0.55555555555556f, w2 = (double)0.88888888888889f;
    if (i == 1) { *a = -a3; *w = w1; }
    else if (i == 2) { *a = 0.0; *w = w2; }
    else { *a = a3; *w = w1; }
}

/* Femlib/hexa.f:906-1050 (HEXA35 -> PRIMAS with LINHEX, 3x3x3 points, GC=-1):
 * consistent mass EM(24,24) = sum_gp w*rho*det(J)/GC * N^T N per direction. */
int hex8_mass(const double X[8], const double Y[8], const double Z[8], double rho, double EM[576])
{
    memset(EM, 0, 576 * sizeof(double));
    const double GC = -1.0;
    double xp[3][8];
    for (int j = 0; j < 8; j++) { xp[0][j] = X[j]; xp[1][j] = Y[j]; xp[2][j] = Z[j]; }
    for (int iz = 1; iz <= 3; iz++) {
        double zeta, zw; gaqdta3(iz, &zeta, &zw);
        for (int iy = 1; iy <= 3; iy++) {
            double eta, yw; gaqdta3(iy, &eta, &yw);
            for (int ix = 1; ix <= 3; ix++) {
                double xsi, xw; gaqdta3(ix, &xsi, &xw);
                double sh[3][8], aj[3][3], N[8];
                for (int j = 0; j < 8; j++) {
                    double d[3]; linhex(j, xsi, eta, zeta, d, 1);
                    sh[0][j] = d[0]; sh[1][j] = d[1]; sh[2][j] = d[2];
                }
                /* ISOJAC: AJAC = SH * XP^T (3x3), AJAC(i,j) = sum_J SH(i,J)*XP(j,J) */
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        double s = 0.0;
                        for (int g = 0; g < 8; g++) s += sh[i][g] * xp[j][g];
                        aj[i][j] = s;
                    }
                for (int j = 0; j < 8; j++) linhex(j, xsi, eta, zeta, &N[j], 0);
                /* DET3(AJAC) with AJAC column-major J(i,j) = aj[i][j] */
                double f1 = aj[1][1] * aj[2][2] - aj[1][2] * aj[2][1];
                double f2 = aj[1][2] * aj[2][0] - aj[1][0] * aj[2][2];
                double f3 = aj[1][0] * aj[2][1] - aj[1][1] * aj[2][0];
                double det = aj[0][0] * f1 + aj[0][1] * f2 + aj[0][2] * f3;
                double c = xw * yw * zw * rho * det / GC;
                for (int a = 0; a < 8; a++)
                    for (int b = 0; b < 8; b++) {
                        double v = c * N[a] * N[b];
                        for (int k = 0; k < 3; k++) EM[(3 * a + k) + 24 * (3 * b + k)] += v;
                    }
            }
        }
    }
    return 0;
}

/* Femlib -> inadd.f90:793-808 (lumpSolidMass): row sums and element mass. */
double lump_solid_mass(int n, const double *MC, double *ML)
{
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += MC[j + n * i];
        ML[i] = s;
    }
    double m = 0.0;
    for (int i = 0; i < n; i += 3) m += ML[i];
    return m;
}
