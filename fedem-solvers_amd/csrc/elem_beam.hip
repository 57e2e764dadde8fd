// gfx950 kernel of the 2-noded 12-DOF beam (ADD11, vpmReducer/inadd.f90:
// 829-930): BEAM31 stiffness (BELS31 local Timoshenko matrix, DCOS30 axes,
// MPRO30 congruence; Femlib/beam.f, beamaux.f) and the BEAM35 consistent
// mass, for straight beams with nodes on the neutral axis and no pin flags.
//
// One lane per (element, node-pair block): lanes 3e, 3e+1, 3e+2 write the
// 6x6 blocks (1,1), (2,1), (2,2) of element e, so a wave covers 21 elements
// and every lane does the same short closed-form work.
//
// Property record (NPROP doubles): E, nu, rho, G, A, I1, I2, J, v(3), K1, K2,
// NSM with the Nastran PBAR/CBAR meaning.  I1 bends in the local x-y plane
// (the plane of the orientation vector v), the shear factors of BELS31 are
// kappa = 1/K (K = 0: no shear deformation), the density carries NSM/A and
// the rotational mass uses I1 + I2.
#include "kernels.hpp"

namespace fgpu {
namespace {

// symmetric 12x12 local stiffness entry (1-based), BELS31 (beam.f:680-715)
struct BeamK {
  double k11, k22, k33, k44, k35, k26, k55, k66, k511, k612;
  __device__ double at(int r, int c) const {
    if (r > c) { int t = r; r = c; c = t; }
    switch (r * 16 + c) {
      case 1 * 16 + 1: case 7 * 16 + 7: return k11;
      case 1 * 16 + 7: return -k11;
      case 2 * 16 + 2: case 8 * 16 + 8: return k22;
      case 2 * 16 + 8: return -k22;
      case 3 * 16 + 3: case 9 * 16 + 9: return k33;
      case 3 * 16 + 9: return -k33;
      case 4 * 16 + 4: case 10 * 16 + 10: return k44;
      case 4 * 16 + 10: return -k44;
      case 3 * 16 + 5: case 3 * 16 + 11: return k35;
      case 5 * 16 + 9: case 9 * 16 + 11: return -k35;
      case 2 * 16 + 6: case 2 * 16 + 12: return k26;
      case 6 * 16 + 8: case 8 * 16 + 12: return -k26;
      case 5 * 16 + 5: case 11 * 16 + 11: return k55;
      case 6 * 16 + 6: case 12 * 16 + 12: return k66;
      case 5 * 16 + 11: return k511;
      case 6 * 16 + 12: return k612;
      default: return 0.0;
    }
  }
};

// symmetric 12x12 local consistent mass entry (1-based), BEAM35 IOP 3 (beam.f:460-485)
struct BeamM {
  double m11, m17, m22, m26, m28, m212, m44, m410, m55, m511;
  __device__ double at(int r, int c) const {
    if (r > c) { int t = r; r = c; c = t; }
    switch (r * 16 + c) {
      case 1 * 16 + 1: case 7 * 16 + 7: return m11;
      case 1 * 16 + 7: return m17;
      case 2 * 16 + 2: case 3 * 16 + 3: case 8 * 16 + 8: case 9 * 16 + 9: return m22;
      case 2 * 16 + 6: case 9 * 16 + 11: return m26;
      case 3 * 16 + 5: case 8 * 16 + 12: return -m26;
      case 2 * 16 + 8: case 3 * 16 + 9: return m28;
      case 2 * 16 + 12: case 5 * 16 + 9: return m212;
      case 3 * 16 + 11: case 6 * 16 + 8: return -m212;
      case 4 * 16 + 4: case 10 * 16 + 10: return m44;
      case 4 * 16 + 10: return m410;
      case 5 * 16 + 5: case 6 * 16 + 6: case 11 * 16 + 11: case 12 * 16 + 12: return m55;
      case 5 * 16 + 11: case 6 * 16 + 12: return m511;
      default: return 0.0;
    }
  }
};

}  // namespace

__global__ __launch_bounds__(256) void k_beam(int nbeam, const int* __restrict__ elist, const int* __restrict__ eptr,
                                              const int* __restrict__ enodes, const double* __restrict__ xyz,
                                              const double* __restrict__ eprop, const int64_t* __restrict__ eoff,
                                              double* __restrict__ Kp, double* __restrict__ Mp,
                                              int* __restrict__ ebad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * nbeam) return;
  const int e = elist[t / 3], blk = t % 3;
  const int na = blk == 0 ? 0 : 1, nb = blk == 2 ? 1 : 0;
  const double* pr = eprop + (int64_t)NPROP * e;
  const int n1 = enodes[eptr[e]], n2 = enodes[eptr[e] + 1];
  const double x1[3] = {xyz[3 * n1], xyz[3 * n1 + 1], xyz[3 * n1 + 2]};
  const double x2[3] = {xyz[3 * n2], xyz[3 * n2 + 1], xyz[3 * n2 + 2]};
  const double v[3] = {pr[P_DX], pr[P_DY], pr[P_DZ]};
  // DCOS30 with the third point node 1 + (e_x x v): rows of C are local axes
  const double ex[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
  const double z[3] = {ex[1] * v[2] - ex[2] * v[1], ex[2] * v[0] - ex[0] * v[2], ex[0] * v[1] - ex[1] * v[0]};
  const double p3[3] = {x1[0] + z[0], x1[1] + z[1], x1[2] + z[2]};
  double C[3][3];
  const double BL = sqrt(ex[0] * ex[0] + ex[1] * ex[1] + ex[2] * ex[2]);
  bool bad = !(BL > 0.0);
  {
    const double ab = BL > 0.0 ? BL : 1.0;
    C[0][0] = ex[0] / ab; C[0][1] = ex[1] / ab; C[0][2] = ex[2] / ab;
    const double d[3] = {p3[0] - x1[0], p3[1] - x1[1], p3[2] - x1[2]};
    double cx = C[0][2] * d[1] - C[0][1] * d[2];
    double cy = C[0][0] * d[2] - C[0][2] * d[0];
    double cz = C[0][1] * d[0] - C[0][0] * d[1];
    double n = sqrt(cx * cx + cy * cy + cz * cz);
    if (!(n > 0.0)) { bad = true; n = 1.0; }
    C[1][0] = cx / n; C[1][1] = cy / n; C[1][2] = cz / n;
    cx = C[0][1] * C[1][2] - C[0][2] * C[1][1];
    cy = C[0][2] * C[1][0] - C[0][0] * C[1][2];
    cz = C[0][0] * C[1][1] - C[0][1] * C[1][0];
    n = sqrt(cx * cx + cy * cy + cz * cz);
    C[2][0] = cx / n; C[2][1] = cy / n; C[2][2] = cz / n;
  }
  const double E = pr[P_E], G = pr[P_G], A = pr[P_A], RIY = pr[P_IY], RIZ = pr[P_IZ], RIT = pr[P_IT];
  if (!(G > 1.0e-16) || !(A > 1.0e-16)) bad = true;
  const double CAY = pr[P_KY] > 0.0 ? 1.0 / pr[P_KY] : 0.0, CAZ = pr[P_KZ] > 0.0 ? 1.0 / pr[P_KZ] : 0.0;
  const double EA = E * A, EIY = E * RIY, EIZ = E * RIZ, GIT = G * RIT;
  const double ALY = 12.0 * CAY * EIY / (A * G * BL * BL), ALZ = 12.0 * CAZ * EIZ / (A * G * BL * BL);
  BeamK k;
  k.k11 = EA / BL;
  k.k22 = 12. * EIY / (BL * BL * BL * (1. + ALY));
  k.k33 = 12. * EIZ / (BL * BL * BL * (1. + ALZ));
  k.k44 = GIT / BL;
  k.k35 = -.5 * BL * k.k33;
  k.k26 = .5 * BL * k.k22;
  k.k55 = EIZ * (4. + ALZ) / (BL * (1. + ALZ));
  k.k66 = EIY * (4. + ALY) / (BL * (1. + ALY));
  k.k511 = EIZ * (2. - ALZ) / (BL * (1. + ALZ));
  k.k612 = EIY * (2. - ALY) / (BL * (1. + ALY));
  const double rho = pr[P_RHO] + (A > 1.0e-16 ? pr[P_NSM] / A : 0.0);
  const double AM = BL * A * rho, AM2 = BL * AM, AM3 = BL * AM2, RM = (RIY + RIZ) * rho;
  BeamM m;
  m.m11 = AM / 3.;
  m.m17 = AM / 6.;
  m.m22 = 13. * AM / 35.;
  m.m26 = 11. * AM2 / 210.;
  m.m28 = 9. * AM / 70.;
  m.m212 = -13. * AM2 / 420.;
  m.m44 = BL * RM / 3.;
  m.m410 = BL * RM / 6.;
  m.m55 = AM3 / 105.;
  m.m511 = -AM3 / 140.;
  // global block = C^T L C per 3x3 sub-block (MPRO30)
  const int64_t off = eoff[e] + 36 * blk;
  for (int sa = 0; sa < 2; sa++)
    for (int sb = 0; sb < 2; sb++) {
      const int r0 = 6 * na + 3 * sa + 1, c0 = 6 * nb + 3 * sb + 1;
      double LK[3][3], LM[3][3];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          LK[i][j] = k.at(r0 + i, c0 + j);
          LM[i][j] = m.at(r0 + i, c0 + j);
        }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          double sk = 0.0, sm = 0.0;
          for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) {
              const double w = C[p][i] * C[q][j];
              sk += w * LK[p][q];
              sm += w * LM[p][q];
            }
          Kp[off + (3 * sa + i) * 6 + 3 * sb + j] = sk;
          Mp[off + (3 * sa + i) * 6 + 3 * sb + j] = sm;
        }
    }
  if (bad && blk == 0) ebad[e] = -3;
}

void launch_beam(hipStream_t st, int nbeam, const int* elist, const int* eptr, const int* enodes, const double* xyz,
                 const double* eprop, const int64_t* eoff, double* Kp, double* Mp, int* ebad) {
  if (nbeam > 0)
    k_beam<<<(unsigned)((3 * (int64_t)nbeam + 255) / 256), 256, 0, st>>>(nbeam, elist, eptr, enodes, xyz, eprop, eoff,
                                                                         Kp, Mp, ebad);
}

}  // namespace fgpu
