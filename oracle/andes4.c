/*
 * ORACLE — test infrastructure only (see elements.c).
 *
 * Plain-C restatement of the 4-noded ANDES shell (ADD_ANDES4,
 * vpmReducer/inadd.f90:1407-1552): the local stiffness of
 * Femlib/andes4ShellModule.f90 (2x2 Gauss, membrane alphaNR = 1.5 with the
 * higher-order part scaled by sqrt(beta) = sqrt(0.9), bending with quadratic
 * lumping, lumpType 2), the 3x3 block rotation to global axes, the rigid-body
 * projector of Femlib/pmatStiff.f90 (K = P^T K P) and the lumped mass of
 * Femlib/fqs.f FQS35 = qmrf.f QMRF35 (membrane) + qbek.f QBEK35 (bending).
 *
 * Arrays follow the Fortran layout of the reference: a 4x4 difference array
 * d(i,j) = c(i) - c(j) is stored d[(i-1) + 4*(j-1)]; matrices column-major.
 *
 * Parity pins (tests/test_oracle_elements.py): trel, Kloc and P against the
 * oracle/_ref build of the reference's own Femlib routines
 * (tests/golden/femlib_ref_vectors.json), FQS35 likewise; invariants (rigid
 * body modes in the null space of P^T K P, symmetry, total lumped mass).
 */
#include <math.h>
#include <string.h>

void iso_mat2d(double E, double nu, double C[9]);

#define D4(a, i, j) (a)[((i) - 1) + 4 * ((j) - 1)]

/* manipMatrixModule.f90:373-415 (invert33), column-major */
static int invert33(const double a[9], double b[9])
{
#define A(i, j) a[((i) - 1) + 3 * ((j) - 1)]
#define B(i, j) b[((i) - 1) + 3 * ((j) - 1)]
    double det = A(1, 1) * (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)) - A(1, 2) * (A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3)) +
                 A(1, 3) * (A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2));
    if (fabs(det) < 1.0e-16) return -1;
    B(1, 1) = (A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3)) / det;
    B(1, 2) = -(A(1, 2) * A(3, 3) - A(3, 2) * A(1, 3)) / det;
    B(1, 3) = (A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3)) / det;
    B(2, 1) = -(A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3)) / det;
    B(2, 2) = (A(1, 1) * A(3, 3) - A(3, 1) * A(1, 3)) / det;
    B(2, 3) = -(A(1, 1) * A(2, 3) - A(2, 1) * A(1, 3)) / det;
    B(3, 1) = (A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2)) / det;
    B(3, 2) = -(A(1, 1) * A(3, 2) - A(3, 1) * A(1, 2)) / det;
    B(3, 3) = (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) / det;
#undef A
#undef B
    return 0;
}

/* andes4ShellModule.f90:1011-1051 (shape4) */
static void shape4(double xsi, double eta, double nv[4], double nx[4], double ne[4])
{
    if (nv) {
        nv[0] = (1.0 - xsi) * (1.0 - eta) * 0.25;
        nv[1] = (1.0 + xsi) * (1.0 - eta) * 0.25;
        nv[2] = (1.0 + xsi) * (1.0 + eta) * 0.25;
        nv[3] = (1.0 - xsi) * (1.0 + eta) * 0.25;
    }
    if (nx) {
        nx[0] = -(1.0 - eta) * 0.25;
        nx[1] = (1.0 - eta) * 0.25;
        nx[2] = (1.0 + eta) * 0.25;
        nx[3] = -(1.0 + eta) * 0.25;
    }
    if (ne) {
        ne[0] = -(1.0 - xsi) * 0.25;
        ne[1] = -(1.0 + xsi) * 0.25;
        ne[2] = (1.0 + xsi) * 0.25;
        ne[3] = (1.0 - xsi) * 0.25;
    }
}

/* andes4ShellModule.f90:392-438 (jacobi_quad): |g_xsi x g_eta|; xyz[n][c]
 * node n, component c; jac (optional) rows g_xsi, g_eta. */
static double jacobi_quad(double xsi, double eta, const double xyz[4][3], double jac[2][3])
{
    double nx[4], ne[4], gx[3], ge[3];
    shape4(xsi, eta, 0, nx, ne);
    for (int c = 0; c < 3; c++) {
        gx[c] = nx[0] * xyz[0][c] + nx[1] * xyz[1][c] + nx[2] * xyz[2][c] + nx[3] * xyz[3][c];
        ge[c] = ne[0] * xyz[0][c] + ne[1] * xyz[1][c] + ne[2] * xyz[2][c] + ne[3] * xyz[3][c];
    }
    if (jac)
        for (int c = 0; c < 3; c++) { jac[0][c] = gx[c]; jac[1][c] = ge[c]; }
    double n0 = gx[1] * ge[2] - gx[2] * ge[1];
    double n1 = gx[2] * ge[0] - gx[0] * ge[2];
    double n2 = gx[0] * ge[1] - gx[1] * ge[0];
    return sqrt(n0 * n0 + n1 * n1 + n2 * n2);
}

/* andes4ShellModule.f90:342-389 (bmean_quad), 2x2 Gauss mean of nodal
 * values bn[n][r][c] (4 x 3 x ncol) -> mean[r][c] */
static void bmean_quad(int ncol, const double *bn, const double xl[4][3], double *mean)
{
    const double g = sqrt(1.0 / 3.0), pos[2] = {-g, g};
    memset(mean, 0, sizeof(double) * 3 * ncol);
    double area = 0.0;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            double fac = jacobi_quad(pos[i], pos[j], xl, 0) * 1.0 * 1.0, n[4];
            shape4(pos[i], pos[j], n, 0, 0);
            for (int k = 0; k < 4; k++) n[k] *= fac;
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < ncol; c++)
                    mean[r * ncol + c] += n[0] * bn[(0 * 3 + r) * ncol + c] + n[1] * bn[(1 * 3 + r) * ncol + c] +
                                          n[2] * bn[(2 * 3 + r) * ncol + c] + n[3] * bn[(3 * 3 + r) * ncol + c];
            area += fac;
        }
    for (int k = 0; k < 3 * ncol; k++) mean[k] /= area;
}

/* andes4ShellModule.f90:189-290 (bben_nod_quad): nodal bending strains
 * bb[n][r][c], 4 x 3 x 12, columns w(1..4), thx(1..4)... as ordered there */
static int bben_nod_quad(const double *xij, const double *yij, double bb[4][3][12])
{
    static const int stnod[4][3] = {{4, 2, 3}, {1, 3, 4}, {2, 4, 1}, {3, 1, 2}};
    double lij[4][4], sij[4][4][2], nij[4][4][2];
    memset(bb, 0, sizeof(double) * 144);
    for (int i = 1; i <= 4; i++)
        for (int k = 1; k <= 4; k++) {
            int j = 1 + (i + k - 1) % 4;
            if (j == i) continue; /* the reference's diagonal entries are 0/0 and never read */
            double l = sqrt(D4(xij, i, j) * D4(xij, i, j) + D4(yij, i, j) * D4(yij, i, j));
            lij[i - 1][j - 1] = l;
            lij[j - 1][i - 1] = l;
            sij[i - 1][j - 1][0] = D4(xij, j, i) / l;
            sij[i - 1][j - 1][1] = D4(yij, j, i) / l;
            nij[i - 1][j - 1][0] = sij[i - 1][j - 1][1];
            nij[i - 1][j - 1][1] = -sij[i - 1][j - 1][0];
            sij[j - 1][i - 1][0] = -sij[i - 1][j - 1][0];
            sij[j - 1][i - 1][1] = -sij[i - 1][j - 1][1];
            nij[j - 1][i - 1][0] = -nij[i - 1][j - 1][0];
            nij[j - 1][i - 1][1] = -nij[i - 1][j - 1][1];
        }
    for (int in = 1; in <= 4; in++)
        for (int r = 1; r <= 3; r++) {
            int jn = stnod[in - 1][r - 1];
            double l = lij[in - 1][jn - 1];
            double *b = bb[in - 1][r - 1];
            b[in - 1] = -6.0 / (l * l);
            b[in + 3] = -4.0 * nij[in - 1][jn - 1][0] / l;
            b[in + 7] = -4.0 * nij[in - 1][jn - 1][1] / l;
            b[jn - 1] = -b[in - 1];
            b[jn + 3] = b[in + 3] * 0.5;
            b[jn + 7] = b[in + 7] * 0.5;
        }
    for (int in = 1; in <= 4; in++) {
        double tinv[9], tran[9];
        for (int r = 1; r <= 3; r++) {
            int jn = stnod[in - 1][r - 1];
            const double *s = sij[in - 1][jn - 1];
            tinv[(r - 1) + 0] = s[0] * s[0];
            tinv[(r - 1) + 3] = s[1] * s[1];
            tinv[(r - 1) + 6] = s[0] * s[1];
        }
        if (invert33(tinv, tran)) return -1;
        for (int c = 0; c < 12; c++) {
            double v[3];
            for (int r = 0; r < 3; r++)
                v[r] = tran[r] * bb[in - 1][0][c] + tran[r + 3] * bb[in - 1][1][c] + tran[r + 6] * bb[in - 1][2][c];
            for (int r = 0; r < 3; r++) bb[in - 1][r][c] = v[r];
        }
    }
    return 0;
}

static double cross2d(const double a[2], const double b[2]) { return a[0] * b[1] - b[0] * a[1]; }

/* andes4ShellModule.f90:479-750 (bmqh_nod_quad): nodal higher-order
 * membrane strains bm[n][r][c], 4 x 3 x 7 */
static int bmqh_nod_quad(const double *xij, const double *yij, double bm[4][3][7])
{
    const double rho1 = 0.1, rho2 = -rho1, rho3 = rho2, rho4 = rho1, alpha = 0.1;
    const double beta1 = 0.5 + rho1, beta2 = 0.0, rho5 = 0.0, rho6 = beta1 - rho1, rho7 = 0.0, rho8 = -rho6;
    double rc1 = (D4(xij, 2, 1) + D4(xij, 3, 1) + D4(xij, 4, 1)) * 0.25;
    double rc2 = (D4(yij, 2, 1) + D4(yij, 3, 1) + D4(yij, 4, 1)) * 0.25;
    double rv[4][2];
    for (int n = 1; n <= 4; n++) {
        rv[n - 1][0] = D4(xij, n, 1) - rc1;
        rv[n - 1][1] = D4(yij, n, 1) - rc2;
    }
    double Xi[2], Eta[2], len;
    Xi[0] = (rv[1][0] + rv[2][0]) - (rv[0][0] + rv[3][0]);
    Xi[1] = (rv[1][1] + rv[2][1]) - (rv[0][1] + rv[3][1]);
    len = sqrt(Xi[0] * Xi[0] + Xi[1] * Xi[1]);
    Xi[0] /= len; Xi[1] /= len;
    double lXi = len * 0.5;
    Eta[0] = (rv[2][0] + rv[3][0]) - (rv[0][0] + rv[1][0]);
    Eta[1] = (rv[2][1] + rv[3][1]) - (rv[0][1] + rv[1][1]);
    len = sqrt(Eta[0] * Eta[0] + Eta[1] * Eta[1]);
    Eta[0] /= len; Eta[1] /= len;
    double lEta = len * 0.5;
    double dv13[2] = {D4(xij, 3, 1), D4(yij, 3, 1)};
    double l13 = sqrt(dv13[0] * dv13[0] + dv13[1] * dv13[1]);
    double u13[2] = {dv13[0] / l13, dv13[1] / l13};
    double dv24[2] = {D4(xij, 4, 2), D4(yij, 4, 2)};
    double l24 = sqrt(dv24[0] * dv24[0] + dv24[1] * dv24[1]);
    double u24[2] = {dv24[0] / l24, dv24[1] / l24};
    double fX[4], fE[4];
    for (int i = 0; i < 4; i++) {
        fX[i] = cross2d(Xi, rv[i]) / lXi;
        fE[i] = cross2d(Eta, rv[i]) / lEta;
    }
    double mX = (fX[0] + fX[1] + fX[2] + fX[3]) * 0.25, mE = (fE[0] + fE[1] + fE[2] + fE[3]) * 0.25;
    for (int i = 0; i < 4; i++) { fX[i] -= mX; fE[i] -= mE; }
    fX[0] = -fX[0]; fX[1] = -fX[1];
    fE[1] = -fE[1]; fE[2] = -fE[2];
    mX = (fX[0] + fX[1] + fX[2] + fX[3]) * 0.25;
    mE = (fE[0] + fE[1] + fE[2] + fE[3]) * 0.25;
    double f13 = cross2d(u13, dv24) * 0.5 / l13;
    double f24 = -cross2d(u24, dv13) * 0.5 / l24;
    double d13x = u13[0] * Xi[0] + u13[1] * Xi[1], d13e = u13[0] * Eta[0] + u13[1] * Eta[1];
    double d24x = u24[0] * Xi[0] + u24[1] * Xi[1], d24e = u24[0] * Eta[0] + u24[1] * Eta[1];
    /* rows: xi-strain, eta-strain, diagonal strain; node order per the table
     * of andes4ShellModule.f90:610-690 */
    const double aE = alpha * lEta / lXi, aX = alpha * lXi / lEta;
    double t[4][3][7] = {
        {{rho1 * fX[0], rho2 * fX[0], rho3 * fX[0], rho4 * fX[0], aE, -beta1 * fX[0] / (mX * lXi), 0.0},
         {-rho1 * fE[0], -rho4 * fE[0], -rho3 * fE[0], -rho2 * fE[0], -aX, 0.0, -beta1 * fE[0] / (mE * lEta)},
         {rho5 * f24, rho6 * f24, rho7 * f24, rho8 * f24, 0.0, beta2 * d24x / l24, -beta2 * d24e / l24}},
        {{-rho2 * fX[1], -rho1 * fX[1], -rho4 * fX[1], -rho3 * fX[1], -aE, -beta1 * fX[1] / (mX * lXi), 0.0},
         {rho4 * fE[1], rho1 * fE[1], rho2 * fE[1], rho3 * fE[1], aX, 0.0, beta1 * fE[1] / (mE * lEta)},
         {rho8 * f13, rho5 * f13, rho6 * f13, rho7 * f13, 0.0, -beta2 * d13x / l13, beta2 * d13e / l13}},
        {{rho3 * fX[2], rho4 * fX[2], rho1 * fX[2], rho2 * fX[2], aE, beta1 * fX[2] / (mX * lXi), 0.0},
         {-rho3 * fE[2], -rho2 * fE[2], -rho1 * fE[2], -rho4 * fE[2], -aX, 0.0, beta1 * fE[2] / (mE * lEta)},
         {rho7 * f24, rho8 * f24, rho5 * f24, rho6 * f24, 0.0, -beta2 * d24x / l24, beta2 * d24e / l24}},
        {{-rho4 * fX[3], -rho3 * fX[3], -rho2 * fX[3], -rho1 * fX[3], -aE, beta1 * fX[3] / (mX * lXi), 0.0},
         {rho2 * fE[3], rho3 * fE[3], rho4 * fE[3], rho1 * fE[3], aX, 0.0, -beta1 * fE[3] / (mE * lEta)},
         {rho6 * f13, rho7 * f13, rho8 * f13, rho5 * f13, 0.0, beta2 * d13x / l13, -beta2 * d13e / l13}}};
    memcpy(bm, t, sizeof t);
    double ti13[9], ti24[9], tr13[9], tr24[9];
    ti24[0] = Xi[0] * Xi[0]; ti24[3] = Xi[1] * Xi[1]; ti24[6] = Xi[0] * Xi[1];
    ti24[1] = Eta[0] * Eta[0]; ti24[4] = Eta[1] * Eta[1]; ti24[7] = Eta[0] * Eta[1];
    memcpy(ti13, ti24, sizeof ti13);
    ti13[2] = u24[0] * u24[0]; ti13[5] = u24[1] * u24[1]; ti13[8] = u24[0] * u24[1];
    ti24[2] = u13[0] * u13[0]; ti24[5] = u13[1] * u13[1]; ti24[8] = u13[0] * u13[1];
    if (invert33(ti13, tr13) || invert33(ti24, tr24)) return -1;
    for (int n = 0; n < 4; n++) {
        const double *tr = (n % 2 == 0) ? tr13 : tr24; /* nodes 1,3 and 2,4 */
        for (int c = 0; c < 7; c++) {
            double v[3];
            for (int r = 0; r < 3; r++) v[r] = tr[r] * bm[n][0][c] + tr[r + 3] * bm[n][1][c] + tr[r + 6] * bm[n][2][c];
            for (int r = 0; r < 3; r++) bm[n][r][c] = v[r];
        }
    }
    return 0;
}

/* andes4ShellModule.f90:753-863 (hmqv_quad): hm[7][12] */
static int hmqv_quad(const double *xij, const double *yij, const double *zij, double hm[7][12])
{
    double rc1 = (D4(xij, 2, 1) + D4(xij, 3, 1) + D4(xij, 4, 1)) * 0.25;
    double rc2 = (D4(yij, 2, 1) + D4(yij, 3, 1) + D4(yij, 4, 1)) * 0.25;
    double rc3 = (D4(zij, 2, 1) + D4(zij, 3, 1) + D4(zij, 4, 1)) * 0.25;
    double rv[4][3];
    for (int n = 1; n <= 4; n++) {
        rv[n - 1][0] = D4(xij, n, 1) - rc1;
        rv[n - 1][1] = D4(yij, n, 1) - rc2;
        rv[n - 1][2] = D4(zij, n, 1) - rc3;
    }
    double Xi1 = (rv[1][0] + rv[2][0]) - (rv[0][0] + rv[3][0]);
    double Xi2 = (rv[1][1] + rv[2][1]) - (rv[0][1] + rv[3][1]);
    double fac = sqrt(Xi1 * Xi1 + Xi2 * Xi2);
    Xi1 /= fac; Xi2 /= fac;
    double Eta1 = (rv[2][0] + rv[3][0]) - (rv[0][0] + rv[1][0]);
    double Eta2 = (rv[2][1] + rv[3][1]) - (rv[0][1] + rv[1][1]);
    fac = sqrt(Eta1 * Eta1 + Eta2 * Eta2);
    Eta1 /= fac; Eta2 /= fac;
    double jac[2][3];
    fac = 0.0625 / jacobi_quad(0.0, 0.0, rv, jac);
    memset(hm, 0, sizeof(double) * 84);
    for (int i = 0; i < 4; i++) {
        for (int c = 8; c < 12; c++) hm[i][c] = -0.25;
        hm[i][i + 8] = 0.75;
    }
    hm[4][0] = -D4(xij, 2, 4) * fac; hm[4][1] = -D4(xij, 3, 1) * fac;
    hm[4][2] = -D4(xij, 4, 2) * fac; hm[4][3] = -D4(xij, 1, 3) * fac;
    hm[4][4] = -D4(yij, 2, 4) * fac; hm[4][5] = -D4(yij, 3, 1) * fac;
    hm[4][6] = -D4(yij, 4, 2) * fac; hm[4][7] = -D4(yij, 1, 3) * fac;
    for (int c = 8; c < 12; c++) hm[4][c] = 0.25;
    /* tnat = inverse of the transposed 2x2 in-plane Jacobian (invert22) */
    double a11 = jac[0][0], a12 = jac[1][0], a21 = jac[0][1], a22 = jac[1][1];
    double det = a11 * a22 - a21 * a12;
    if (fabs(det) < 1.0e-16) return -1;
    double t11 = a22 / det, t22 = a11 / det, t12 = -a12 / det, t21 = -a21 / det;
    double rn[4][2];
    for (int n = 0; n < 4; n++) {
        rn[n][0] = t11 * rv[n][0] + t12 * rv[n][1];
        rn[n][1] = t21 * rv[n][0] + t22 * rv[n][1];
    }
    fac = (rn[0][0] * rn[0][1] + rn[1][0] * rn[1][1] + rn[2][0] * rn[2][1] + rn[3][0] * rn[3][1]) * 0.25;
    static const int nodeof[4] = {2, 3, 0, 1}; /* columns 1..4 take nodes 3,4,1,2 */
    for (int c = 0; c < 4; c++) {
        double q = rn[nodeof[c]][0] * rn[nodeof[c]][1] - fac;
        hm[5][c] = q * Xi1; hm[5][c + 4] = q * Xi2;
        hm[6][c] = q * Eta1; hm[6][c + 4] = q * Eta2;
    }
    return 0;
}

/* andes4ShellModule.f90:117-186 (lump4ben, ltype 2) and 441-476
 * (lump4mem): Lt[r][c], c = node + 4*(k-1) */
static void lump4ben(const double *x, const double *y, double Lt[3][12])
{
    for (int i = 1; i <= 4; i++) {
        int j = 1 + i % 4, l = 1 + (i + 2) % 4;
        double len2 = D4(x, i, j) * D4(x, i, j) + D4(y, i, j) * D4(y, i, j);
        double ccij = D4(x, i, j) * D4(x, i, j) / len2, ssij = D4(y, i, j) * D4(y, i, j) / len2;
        double csij = D4(x, i, j) * D4(y, i, j) / len2;
        len2 = D4(x, l, i) * D4(x, l, i) + D4(y, l, i) * D4(y, l, i);
        double ccli = D4(x, l, i) * D4(x, l, i) / len2, ssli = D4(y, l, i) * D4(y, l, i) / len2;
        double csli = D4(x, l, i) * D4(y, l, i) / len2;
        int c1 = i - 1, c2 = i + 3, c3 = i + 7;
        Lt[0][c1] = (ssij * D4(x, i, j) + ssli * D4(x, l, i)) * 0.5;
        Lt[1][c1] = (ccij * D4(x, i, j) + ccli * D4(x, l, i)) * 0.5;
        Lt[2][c1] = -ccij * D4(y, i, j) - ccli * D4(y, l, i);
        Lt[0][c2] = (ssij * D4(y, i, j) + ssli * D4(y, l, i)) * 0.5;
        Lt[1][c2] = (ccij * D4(y, i, j) + ccli * D4(y, l, i)) * 0.5;
        Lt[2][c2] = -ssij * D4(x, i, j) - ssli * D4(x, l, i);
        Lt[0][c3] = -csij + csli;
        Lt[1][c3] = csij - csli;
        Lt[2][c3] = (ssli - ccli) - (ssij - ccij);
    }
}

static void lump4mem(const double *x, const double *y, double alpha, double Lt[3][12])
{
    for (int i = 1; i <= 4; i++) {
        int j = 1 + i % 4, l = 1 + (i + 2) % 4;
        int c1 = i - 1, c2 = i + 3, c3 = i + 7;
        Lt[0][c1] = D4(y, j, l) * 0.5;
        Lt[1][c1] = 0.0;
        Lt[2][c1] = -D4(x, j, l) * 0.5;
        Lt[0][c2] = 0.0;
        Lt[1][c2] = -D4(x, j, l) * 0.5;
        Lt[2][c2] = D4(y, j, l) * 0.5;
        Lt[0][c3] = (D4(y, l, i) * D4(y, l, i) - D4(y, j, i) * D4(y, j, i)) * alpha / 12.0;
        Lt[1][c3] = (D4(x, l, i) * D4(x, l, i) - D4(x, j, i) * D4(x, j, i)) * alpha / 12.0;
        Lt[2][c3] = (D4(x, j, i) * D4(y, j, i) - D4(x, l, i) * D4(y, l, i)) * alpha / 6.0;
    }
}

/* andes4ShellModule.f90:917-967 (Bmatrix) with StrainDispMatrix4mem/ben
 * (27-114), bh4ben (293-339), bh4mem (866-914): B[r][c], 6 x 24 */
static int andes4_bmatrix(const double *x, const double *y, const double *z, double alphaH, double alphaNR,
                          double xsi, double eta, double B[6][24])
{
    static const int dofm[12] = {1, 7, 13, 19, 2, 8, 14, 20, 6, 12, 18, 24};
    static const int dofb[12] = {4, 10, 16, 22, 5, 11, 17, 23, 3, 9, 15, 21};
    double a = 0.5 * (D4(x, 2, 1) * D4(y, 3, 1) - D4(x, 3, 1) * D4(y, 2, 1) + D4(x, 4, 3) * D4(y, 1, 3) -
                      D4(x, 1, 3) * D4(y, 4, 3));
    double xl[4][3], nv[4];
    for (int n = 1; n <= 4; n++) {
        xl[n - 1][0] = D4(x, n, 1);
        xl[n - 1][1] = D4(y, n, 1);
        xl[n - 1][2] = D4(z, n, 1);
    }
    shape4(xsi, eta, nv, 0, 0);
    /* membrane */
    double Ltm[3][12], bm[4][3][7], bmean[3][7], bq[3][7], hm[7][12], Bm[3][12];
    lump4mem(x, y, alphaNR, Ltm);
    if (bmqh_nod_quad(x, y, bm)) return -1;
    bmean_quad(7, &bm[0][0][0], xl, &bmean[0][0]);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 7; c++)
            bq[r][c] = (nv[0] * bm[0][r][c] + nv[1] * bm[1][r][c] + nv[2] * bm[2][r][c] + nv[3] * bm[3][r][c]) -
                       bmean[r][c];
    if (hmqv_quad(x, y, z, hm)) return -1;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 12; c++) {
            double s = 0.0;
            for (int k = 0; k < 7; k++) s += bq[r][k] * hm[k][c];
            Bm[r][c] = Ltm[r][c] / a + s * alphaH;
        }
    /* bending */
    double Ltb[3][12], bb[4][3][12], bbm[3][12], Bh[3][12], Bb[3][12];
    lump4ben(x, y, Ltb);
    if (bben_nod_quad(x, y, bb)) return -1;
    bmean_quad(12, &bb[0][0][0], xl, &bbm[0][0]);
    for (int r = 0; r < 3; r++)
        for (int j = 1; j <= 12; j++) {
            int k = 1 + (7 + j) % 12;
            Bh[r][k - 1] = (nv[0] * bb[0][r][j - 1] + nv[1] * bb[1][r][j - 1] + nv[2] * bb[2][r][j - 1] +
                            nv[3] * bb[3][r][j - 1]) -
                           bbm[r][j - 1];
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 12; c++) Bb[r][c] = Ltb[r][c] / a + Bh[r][c] * alphaH;
    memset(B, 0, sizeof(double) * 144);
    for (int j = 0; j < 12; j++)
        for (int r = 0; r < 3; r++) {
            B[r][dofm[j] - 1] = Bm[r][j];
            B[r + 3][dofb[j] - 1] = Bb[r][j];
        }
    return 0;
}

/* andes4ShellModule.f90:1054-1103 (Andes4shell_stiffmat): local 24x24
 * stiffness, column-major, of the element at local coordinates X,Y,Z with
 * the 6x6 constitutive matrix Cmat (column-major). */
int andes4_local_stiff(const double X[4], const double Y[4], const double Z[4], const double Cmat[36], double alpha,
                       double beta, double K[576])
{
    const double g = sqrt(1.0 / 3.0), xi[2] = {-g, g};
    double dx[16], dy[16], dz[16], xl[4][3];
    for (int i = 1; i <= 4; i++) {
        xl[i - 1][0] = X[i - 1];
        xl[i - 1][1] = Y[i - 1];
        xl[i - 1][2] = Z[i - 1];
        for (int j = 1; j <= 4; j++) {
            D4(dx, i, j) = X[i - 1] - X[j - 1];
            D4(dy, i, j) = Y[i - 1] - Y[j - 1];
            D4(dz, i, j) = Z[i - 1] - Z[j - 1];
        }
    }
    memset(K, 0, 576 * sizeof(double));
    double alpH = sqrt(beta);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            double wJ = jacobi_quad(xi[i], xi[j], xl, 0) * 1.0 * 1.0, B[6][24], CB[6][24];
            if (andes4_bmatrix(dx, dy, dz, alpH, alpha, xi[i], xi[j], B)) return -1;
            for (int r = 0; r < 6; r++)
                for (int c = 0; c < 24; c++) {
                    double s = 0.0;
                    for (int k = 0; k < 6; k++) s += Cmat[r + 6 * k] * wJ * B[k][c];
                    CB[r][c] = s;
                }
            for (int a = 0; a < 24; a++)
                for (int b = 0; b < 24; b++) {
                    double s = 0.0;
                    for (int k = 0; k < 6; k++) s += B[k][a] * CB[k][b];
                    K[a + 24 * b] += s;
                }
        }
    return 0;
}

/* manipMatrixModule.f90:447-520 (trans3P): tr(3,4) column-major */
int andes4_trans3p(const double P1[3], const double P2[3], const double P3[3], double tr[12])
{
    double *t1 = tr, *t2 = tr + 3, *t3 = tr + 6;
    for (int k = 0; k < 3; k++) {
        t1[k] = P2[k] - P1[k];
        t2[k] = P3[k] - P1[k];
        tr[9 + k] = P1[k];
    }
    t3[0] = t1[1] * t2[2] - t1[2] * t2[1];
    t3[1] = t1[2] * t2[0] - t1[0] * t2[2];
    t3[2] = t1[0] * t2[1] - t1[1] * t2[0];
    double v = t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2];
    if (!(v > 1.0e-16)) return -1;
    v = sqrt(v);
    for (int k = 0; k < 3; k++) t1[k] /= v;
    v = t3[0] * t3[0] + t3[1] * t3[1] + t3[2] * t3[2];
    if (!(v > 1.0e-16)) return -2;
    v = sqrt(v);
    for (int k = 0; k < 3; k++) t3[k] /= v;
    t2[0] = t3[1] * t1[2] - t3[2] * t1[1];
    t2[1] = t3[2] * t1[0] - t3[0] * t1[2];
    t2[2] = t3[0] * t1[1] - t3[1] * t1[0];
    return 0;
}

/* pmatStiff.f90:23-127 (pmat_stiff1): 6*nnod square projector, column-major;
 * coor[3*n + c] */
int pmat_stiff(int nnod, const double *coor, double *P)
{
    const int nd = 6 * nnod;
    double R[24 * 6], cr[3 * 4];
    if (nnod > 4) return -1;
    for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int n = 0; n < nnod; n++) s += coor[3 * n + c];
        s /= (double)nnod;
        for (int n = 0; n < nnod; n++) cr[3 * n + c] = coor[3 * n + c] - s;
    }
    memset(R, 0, sizeof R);
#define RM(i, j) R[(i) - 1 + nd * ((j) - 1)]
    for (int i = 1; i <= nnod; i++) {
        int j = (i - 1) * 6;
        const double *x = cr + 3 * (i - 1);
        RM(j + 1, 1) = 1.0; RM(j + 2, 2) = 1.0; RM(j + 3, 3) = 1.0;
        RM(j + 2, 4) = -x[2]; RM(j + 3, 4) = x[1]; RM(j + 4, 4) = 1.0;
        RM(j + 1, 5) = x[2]; RM(j + 3, 5) = -x[0]; RM(j + 5, 5) = 1.0;
        RM(j + 1, 6) = -x[1]; RM(j + 2, 6) = x[0]; RM(j + 6, 6) = 1.0;
    }
    for (int c = 1; c <= 6; c++) {
        double s = 0.0;
        for (int i = 1; i <= nd; i++) s += RM(i, c) * RM(i, c);
        s = sqrt(s);
        for (int i = 1; i <= nd; i++) RM(i, c) /= s;
    }
    double sub[9], si[9];
    for (int i = 1; i <= 3; i++)
        for (int j = 1; j <= 3; j++) {
            double s = 0.0;
            for (int k = 1; k <= nd; k++) s += RM(k, i + 3) * RM(k, j + 3);
            sub[(i - 1) + 3 * (j - 1)] = s;
        }
    if (invert33(sub, si)) return -1;
    for (int i = 1; i <= nd; i++) {
        double rs[6];
        rs[0] = RM(i, 1); rs[1] = RM(i, 2); rs[2] = RM(i, 3);
        for (int j = 1; j <= 3; j++)
            rs[2 + j] = RM(i, 4) * si[0 + 3 * (j - 1)] + RM(i, 5) * si[1 + 3 * (j - 1)] + RM(i, 6) * si[2 + 3 * (j - 1)];
        for (int j = 1; j <= nd; j++) {
            double s = 0.0;
            for (int k = 1; k <= 6; k++) s += rs[k - 1] * RM(j, k);
            P[(i - 1) + nd * (j - 1)] = -s;
        }
        P[(i - 1) + nd * (i - 1)] += 1.0;
    }
#undef RM
    return 0;
}

/* ADD_ANDES4 (inadd.f90:1452-1545), stiffness part: global element
 * stiffness EK (24x24, column-major) of the shell with corner coordinates
 * xyz[3*n + c], thickness t (mean of the four nodal values) and isotropic
 * E, nu. */
int andes4_stiff(const double xyz[12], double E, double nu, double t, double EK[576])
{
    const double *p1 = xyz, *p2 = xyz + 3;
    double mid[3], tr[12], Tel[9], cg[12], Xl[4], Yl[4], Zl[4];
    for (int k = 0; k < 3; k++) mid[k] = 0.5 * (xyz[6 + k] + xyz[9 + k]);
    if (andes4_trans3p(p1, p2, mid, tr)) return -1;
    /* TelT = trel(1:3,1:3) (columns = local axes); Tel = transpose */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Tel[i + 3 * j] = tr[j + 3 * i];
    for (int n = 0; n < 4; n++)
        for (int c = 0; c < 3; c++) cg[3 * n + c] = xyz[3 * n + c] - xyz[c];
    for (int n = 0; n < 4; n++) {
        const double *v = cg + 3 * n;
        Xl[n] = Tel[0] * v[0] + Tel[3] * v[1] + Tel[6] * v[2];
        Yl[n] = Tel[1] * v[0] + Tel[4] * v[1] + Tel[7] * v[2];
        Zl[n] = Tel[2] * v[0] + Tel[5] * v[1] + Tel[8] * v[2];
    }
    double C3[9], Cm[36];
    iso_mat2d(E, nu, C3);
    double fBen = t * t * t / 12.0;
    memset(Cm, 0, sizeof Cm);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Cm[i + 6 * j] = C3[i + 3 * j] * t;
            Cm[(i + 3) + 6 * (j + 3)] = C3[i + 3 * j] * fBen;
        }
    double K[576], P[576], Kg[576], T1[576];
    if (andes4_local_stiff(Xl, Yl, Zl, Cm, 1.5, 0.9, K)) return -2;
    /* 3x3 blocks to global: TelT * K_ab * Tel */
    for (int a = 0; a < 8; a++)
        for (int b = 0; b < 8; b++) {
            double blk[9], tmp[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 3; k++) s += K[(3 * a + i) + 24 * (3 * b + k)] * Tel[k + 3 * j];
                    tmp[i + 3 * j] = s;
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double s = 0.0;
                    for (int k = 0; k < 3; k++) s += tr[i + 3 * k] * tmp[k + 3 * j];
                    blk[i + 3 * j] = s;
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) Kg[(3 * a + i) + 24 * (3 * b + j)] = blk[i + 3 * j];
        }
    if (pmat_stiff(4, cg, P)) return -3;
    /* EK = P^T Kg P */
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) {
            double s = 0.0;
            for (int k = 0; k < 24; k++) s += Kg[i + 24 * k] * P[k + 24 * j];
            T1[i + 24 * j] = s;
        }
    for (int i = 0; i < 24; i++)
        for (int j = 0; j < 24; j++) {
            double s = 0.0;
            for (int k = 0; k < 24; k++) s += P[k + 24 * i] * T1[k + 24 * j];
            EK[i + 24 * j] = s;
        }
    return 0;
}

/* The single-precision literals of qmrf.f / qbek.f: 2./3., 1./3. and 1./20.
 * are REAL*4 constants promoted to double. */
#define F23 ((double)(2.0f / 3.0f))
#define F13 ((double)(1.0f / 3.0f))
#define F120 ((double)(1.0f / 20.0f))

/* Common triangle part of qmrf.f:1527-1640 and qbek.f:1729-1845: mass m of
 * the triangle (1, I+1, I+2) and the absolute global inertias IX, IY, IZ.
 * Returns 0 for a zero-area triangle. */
static int tri_inertia(const double *X, const double *Y, const double *Z, int I, double thk, double rho, double *m,
                       double *ix, double *iy, double *iz)
{
    double X2 = X[I] - X[0], Y2 = Y[I] - Y[0], Z2 = Z[I] - Z[0];
    double X3 = X[I + 1] - X[0], Y3 = Y[I + 1] - Y[0], Z3 = Z[I + 1] - Z[0];
    double SL21 = sqrt(X2 * X2 + Y2 * Y2 + Z2 * Z2), SL31 = sqrt(X3 * X3 + Y3 * Y3 + Z3 * Z3);
    double COSG = (X3 * X2 + Y3 * Y2 + Z3 * Z2) / (SL31 * SL21);
    double AUX1 = 1. - COSG * COSG, SING = AUX1 > 0 ? sqrt(AUX1) : 0.0;
    double XL2 = SL21, XL3 = SL31 * COSG, YL3 = SL31 * SING;
    double AREA = (XL2 * YL3) / 2.;
    if (!(AREA > 0)) return 0;
    double XTP = ((XL3 + ((XL2 - XL3) / 2.))) * F23;
    double H = YL3, B = XL2, B1 = XL3, B2 = XL2 - XL3;
    double S1 = sqrt((F23 * B1) * (F23 * B1) + (F13 * H) * (F13 * H));
    double S2 = sqrt((F23 * B2) * (F23 * B2) + (F13 * H) * (F13 * H));
    double IXX = (rho * thk * B * H) * (((H * H) / 36.) + ((thk * thk) / 24.));
    double IYY1 = (rho * thk * B1 * H) * (((B1 * B1) / 36.) + ((thk * thk) / 24.));
    double d1 = XTP - F23 * B1;
    IYY1 = IYY1 + ((d1 * d1) * (((B1 * H) / 2.) * rho * thk));
    double IYY2 = (rho * thk * B2 * H) * (((B2 * B2) / 36.) + ((thk * thk) / 24.));
    double d2 = B1 - XTP + F13 * B2;
    IYY2 = IYY2 + ((d2 * d2) * (((B2 * H) / 2.) * rho * thk));
    double IYY = IYY1 + IYY2;
    double IZZ1 = (rho * thk * H * B1) * (((B1 * B1) / 4.) + ((H * H) / 12.));
    IZZ1 = IZZ1 - ((S1 * S1) * rho * thk * (B1 * H) / 2.);
    IZZ1 = IZZ1 + ((d1 * d1) * ((B1 * H) / 2.) * rho * thk);
    double IZZ2 = (rho * thk * H * B2) * (((B2 * B2) / 4.) + ((H * H) / 12.));
    IZZ2 = IZZ2 - ((S2 * S2) * rho * thk * (B2 * H) / 2.);
    double d3 = B1 + F13 * B2 - XTP;
    IZZ2 = IZZ2 + ((d3 * d3) * ((B2 * H) / 2.) * rho * thk);
    double IZZ = IZZ1 + IZZ2;
    double T[3][3];
    T[0][0] = X2; T[1][0] = Y2; T[2][0] = Z2;
    double SL = sqrt(T[0][0] * T[0][0] + T[1][0] * T[1][0] + T[2][0] * T[2][0]);
    T[0][0] /= SL; T[1][0] /= SL; T[2][0] /= SL;
    T[0][1] = X3; T[1][1] = Y3; T[2][1] = Z3;
    T[0][2] = T[1][0] * T[2][1] - T[1][1] * T[2][0];
    T[1][2] = T[0][1] * T[2][0] - T[0][0] * T[2][1];
    T[2][2] = T[0][0] * T[1][1] - T[0][1] * T[1][0];
    SL = sqrt(T[0][2] * T[0][2] + T[1][2] * T[1][2] + T[2][2] * T[2][2]);
    T[0][2] /= SL; T[1][2] /= SL; T[2][2] /= SL;
    T[0][1] = T[1][2] * T[2][0] - T[1][0] * T[2][2];
    T[1][1] = T[0][0] * T[2][2] - T[0][2] * T[2][0];
    T[2][1] = T[0][2] * T[1][0] - T[0][0] * T[1][2];
    *ix = fabs(T[0][0] * IXX + T[0][1] * IYY + T[0][2] * IZZ);
    *iy = fabs(T[1][0] * IXX + T[1][1] * IYY + T[1][2] * IZZ);
    *iz = fabs(T[2][0] * IXX + T[2][1] * IYY + T[2][2] * IZZ);
    *m = (rho * AREA * thk) / 3.;
    return 1;
}

/* fqs.f FQS35 (QMRF35 membrane + QBEK35 bending): lumped diagonal mass
 * EM[24] of the quadrilateral shell, global X,Y,Z, nodal thickness TH. */
void fqs35(const double X[4], const double Y[4], const double Z[4], const double TH[4], double rho, double EM[24])
{
    double thk = (TH[0] + TH[1] + TH[2] + TH[3]) * 0.25;
    double m[2] = {0, 0}, ix[2] = {0, 0}, iy[2] = {0, 0}, iz[2] = {0, 0};
    int ok[2];
    memset(EM, 0, 24 * sizeof(double));
    for (int I = 1; I <= 2; I++) ok[I - 1] = tri_inertia(X, Y, Z, I, thk, rho, &m[I - 1], &ix[I - 1], &iy[I - 1], &iz[I - 1]);
    /* membrane (qmrf.f:1641-1689): translations x, y and the drilling DOF */
    if (ok[0] && ok[1]) {
        double g = F120;
        double a1 = m[0] + m[1], a2 = m[0] + m[1], a6 = g * iz[0] + g * iz[1];
        double a7 = m[0], a8 = m[0], a12 = g * iz[0];
        double a13 = m[0] + m[1], a14 = m[0] + m[1], a18 = g * iz[0] + g * iz[1];
        double a19 = m[1], a20 = m[1], a24 = g * iz[1];
        double e1 = (a1 + a7 + a13 + a19) / 4.0, e2 = (a2 + a8 + a14 + a20) / 4.0, e6 = (a6 + a12 + a18 + a24) / 4.0;
        EM[0] = EM[6] = EM[12] = EM[18] = e1;
        EM[1] = EM[7] = EM[13] = EM[19] = e2;
        EM[5] = EM[11] = EM[17] = EM[23] = e6;
        /* bending (qbek.f:1846-1894): z translation and the two rotations */
        double b3 = m[0] + m[1], b4 = g * ix[0] + g * ix[1], b5 = g * iy[0] + g * iy[1];
        double b9 = m[0], b10 = g * ix[0], b11 = g * iy[0];
        double b15 = m[0] + m[1], b16 = g * ix[0] + g * ix[1], b17 = g * iy[0] + g * iy[1];
        double b21 = m[1], b22 = g * ix[1], b23 = g * iy[1];
        double e3 = (b3 + b9 + b15 + b21) / 4.0, e4 = (b4 + b10 + b16 + b22) / 4.0, e5 = (b5 + b11 + b17 + b23) / 4.0;
        EM[2] = EM[8] = EM[14] = EM[20] = e3;
        EM[3] = EM[9] = EM[15] = EM[21] = e4;
        EM[4] = EM[10] = EM[16] = EM[22] = e5;
    }
}
