/*
 * ORACLE — test infrastructure only (see elements.c).
 *
 * Plain-C restatement of the 2-noded 12-DOF beam of ADD11
 * (vpmReducer/inadd.f90:829-930): stiffness BEAM31 (Femlib/beam.f:11-165)
 * = BELS31 local matrix (beam.f:619-770; Timoshenko shear parameters
 * ALY = 12 kappa_y E I_y / (A G L^2)), DCOS30 direction cosines
 * (beamaux.f:48-104) and the MPRO30 congruence (beamaux.f:106-163); mass
 * BEAM35 (beam.f:396-575), consistent (IOP 3) with the element mass EMASS
 * taken from the diagonal form (IOP 2), as ADD11 does.  Straight beams with
 * nodes on the neutral axis, no pin flags and no cross-section rotation
 * (eccentricities, TRIX30, are zero).
 *
 * Orientation: the third point of DCOS30 lies in the local x-z plane; it is
 * built here as node 1 + (e_x x v) from the Nastran CBAR orientation vector
 * v, which spans the local x-y plane.
 *
 * Parity pins: the CBAR02_ known answer (tests/test_oracle_kat.py), the
 * Euler-Bernoulli / Timoshenko closed forms of a cantilever, rigid-body null
 * space and the consistent-mass totals (tests/test_oracle_elements.py).
 */
#include <math.h>
#include <string.h>

/* DCOS30: rows of C = local axes in global components (global -> local) */
static int dcos30(const double X[3], const double Y[3], const double Z[3], double C[3][3])
{
    double cx = X[1] - X[0], cy = Y[1] - Y[0], cz = Z[1] - Z[0];
    double ab = sqrt(cx * cx + cy * cy + cz * cz);
    if (!(ab > 0)) return -1;
    C[0][0] = cx / ab; C[0][1] = cy / ab; C[0][2] = cz / ab;
    cx = C[0][2] * (Y[2] - Y[0]) - C[0][1] * (Z[2] - Z[0]);
    cy = C[0][0] * (Z[2] - Z[0]) - C[0][2] * (X[2] - X[0]);
    cz = C[0][1] * (X[2] - X[0]) - C[0][0] * (Y[2] - Y[0]);
    ab = sqrt(cx * cx + cy * cy + cz * cz);
    if (!(ab > 0)) return -2;
    C[1][0] = cx / ab; C[1][1] = cy / ab; C[1][2] = cz / ab;
    cx = C[0][1] * C[1][2] - C[0][2] * C[1][1];
    cy = C[0][2] * C[1][0] - C[0][0] * C[1][2];
    cz = C[0][0] * C[1][1] - C[0][1] * C[1][0];
    ab = sqrt(cx * cx + cy * cy + cz * cz);
    C[2][0] = cx / ab; C[2][1] = cy / ab; C[2][2] = cz / ab;
    return 0;
}

/* MPRO30: EK <- T^T EK T with T = diag(C, C, C, C), 12x12 column-major */
static void mpro30(double *EK, const double C[3][3])
{
#define K(i, j) EK[((i) - 1) + 12 * ((j) - 1)]
    for (int I = 0; I <= 9; I += 3)
        for (int J = 0; J <= 9; J += 3) {
            double B[3][3] = {{0}};
            for (int ia = 1; ia <= 3; ia++)
                for (int ka = 1; ka <= 3; ka++)
                    for (int ja = 1; ja <= 3; ja++) B[ja - 1][ia - 1] += C[ka - 1][ja - 1] * K(J + ka, I + ia);
            for (int ia = 1; ia <= 3; ia++)
                for (int ka = 1; ka <= 3; ka++) K(J + ka, I + ia) = 0.0;
            for (int ia = 1; ia <= 3; ia++)
                for (int ka = 1; ka <= 3; ka++)
                    for (int ja = 1; ja <= 3; ja++) K(J + ka, I + ia) += B[ka - 1][ja - 1] * C[ja - 1][ia - 1];
        }
    for (int i = 1; i <= 12; i++)
        for (int j = i; j <= 12; j++) K(i, j) = K(j, i);
#undef K
}

/* third point of DCOS30 from the orientation vector v */
static void beam_points(const double x1[3], const double x2[3], const double v[3], double X[3], double Y[3],
                        double Z[3])
{
    double ex[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
    double z[3] = {ex[1] * v[2] - ex[2] * v[1], ex[2] * v[0] - ex[0] * v[2], ex[0] * v[1] - ex[1] * v[0]};
    X[0] = x1[0]; Y[0] = x1[1]; Z[0] = x1[2];
    X[1] = x2[0]; Y[1] = x2[1]; Z[1] = x2[2];
    X[2] = x1[0] + z[0]; Y[2] = x1[1] + z[1]; Z[2] = x1[2] + z[2];
}

/* BEAM31 + BELS31: global 12x12 stiffness (column-major) */
int beam_stiff(const double x1[3], const double x2[3], const double v[3], double E, double G, double A, double RIY,
               double RIZ, double RIT, double CAY, double CAZ, double EK[144])
{
    double X[3], Y[3], Z[3], C[3][3];
    beam_points(x1, x2, v, X, Y, Z);
    double bx = X[1] - X[0], by = Y[1] - Y[0], bz = Z[1] - Z[0];
    double BL = sqrt(bx * bx + by * by + bz * bz);
    if (!(G > 1.0e-16) || !(A > 1.0e-16)) return -3;
    if (dcos30(X, Y, Z, C)) return -1;
    double EA = E * A, EIY = E * RIY, EIZ = E * RIZ, GIT = G * RIT;
    double ALY = 12.0 * CAY * EIY / (A * G * BL * BL), ALZ = 12.0 * CAZ * EIZ / (A * G * BL * BL);
    memset(EK, 0, 144 * sizeof(double));
#define K(i, j) EK[((i) - 1) + 12 * ((j) - 1)]
    K(1, 1) = EA / BL;
    K(2, 2) = 12. * EIY / (BL * BL * BL * (1. + ALY));
    K(3, 3) = 12. * EIZ / (BL * BL * BL * (1. + ALZ));
    K(4, 4) = GIT / BL;
    K(3, 5) = -.5 * BL * K(3, 3);
    K(2, 6) = .5 * BL * K(2, 2);
    K(5, 5) = EIZ * (4. + ALZ) / (BL * (1. + ALZ));
    K(6, 6) = EIY * (4. + ALY) / (BL * (1. + ALY));
    K(1, 7) = -K(1, 1);
    K(7, 7) = K(1, 1);
    K(2, 8) = -K(2, 2);
    K(6, 8) = -K(2, 6);
    K(8, 8) = K(2, 2);
    K(3, 9) = -K(3, 3);
    K(5, 9) = -K(3, 5);
    K(9, 9) = K(3, 3);
    K(4, 10) = -K(4, 4);
    K(10, 10) = K(4, 4);
    K(3, 11) = K(3, 5);
    K(5, 11) = EIZ * (2. - ALZ) / (BL * (1. + ALZ));
    K(9, 11) = -K(3, 5);
    K(11, 11) = K(5, 5);
    K(2, 12) = K(2, 6);
    K(6, 12) = EIY * (2. - ALY) / (BL * (1. + ALY));
    K(8, 12) = -K(2, 6);
    K(12, 12) = K(6, 6);
    for (int i = 1; i <= 12; i++)
        for (int j = 1; j <= i; j++) K(i, j) = K(j, i);
#undef K
    mpro30(EK, C);
    return 0;
}

/* BEAM35 IOP 3 (consistent, global) and the ADD11 element mass from IOP 2
 * (EM(1,1) + EM(7,1) = rho A L) */
int beam_mass(const double x1[3], const double x2[3], const double v[3], double A, double RHO, double RIX,
              double EM[144], double *emass)
{
    double X[3], Y[3], Z[3], C[3][3];
    beam_points(x1, x2, v, X, Y, Z);
    double bx = X[1] - X[0], by = Y[1] - Y[0], bz = Z[1] - Z[0];
    double BL = sqrt(bx * bx + by * by + bz * bz);
    double AM = BL * A * RHO, HM = AM * 0.5, AM2 = BL * AM, AM3 = BL * AM2, RM = RIX * RHO;
    if (emass) *emass = HM + HM;
    if (dcos30(X, Y, Z, C)) return -1;
    memset(EM, 0, 144 * sizeof(double));
#define M(i, j) EM[((i) - 1) + 12 * ((j) - 1)]
    M(1, 1) = AM / 3.;
    M(1, 7) = AM / 6.;
    M(2, 2) = 13. * AM / 35.;
    M(2, 6) = 11. * AM2 / 210.;
    M(2, 8) = 9. * AM / 70.;
    M(2, 12) = -13. * AM2 / 420.;
    M(3, 3) = M(2, 2);
    M(3, 5) = -M(2, 6);
    M(3, 9) = M(2, 8);
    M(3, 11) = -M(2, 12);
    M(4, 4) = BL * RM / 3.;
    M(4, 10) = BL * RM / 6.;
    M(5, 5) = AM3 / 105.;
    M(5, 9) = M(2, 12);
    M(5, 11) = -AM3 / 140.;
    M(6, 6) = M(5, 5);
    M(6, 8) = M(3, 11);
    M(6, 12) = M(5, 11);
    M(7, 7) = M(1, 1);
    M(8, 8) = M(2, 2);
    M(8, 12) = M(3, 5);
    M(9, 9) = M(8, 8);
    M(9, 11) = M(2, 6);
    M(10, 10) = M(4, 4);
    M(11, 11) = M(5, 5);
    M(12, 12) = M(6, 6);
    for (int i = 1; i <= 12; i++)
        for (int j = i; j <= 12; j++) M(j, i) = M(i, j);
#undef M
    mpro30(EM, C);
    return 0;
}
