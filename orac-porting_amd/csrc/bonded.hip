// bonded.hip -- proper and improper torsions (fptors.f:1-345,
// fitors.f:1-286), harmonic bonds
// (fpbond.f:1-110) and bends with their Urey-Bradley and stretch-stretch
// coupling terms (fpbend.f:1-228) on gfx950.
//
// One thread per term (a torsion is described here; bonds and bends follow
// the same scheme with two and three atoms) computes the dihedral cosine from the three bond
// angles, the energy and the scalar force of its multiplicity, and the
// derivative vectors for its four atoms, and stores the twelve force
// components; a per-atom gather then adds the components of every torsion
// the atom is in, in torsion order, so the force sums are deterministic and
// need no atomics.  FP64 throughout: torsions are few (a few per atom) and
// latency bound.
#include "common.hpp"
#include "kernels.hpp"

namespace orac {

// x**n, integer n >= 1, products left to right (as the compiled Fortran does)
__device__ inline double ipow_ltr(double x, int n) {
  double r = x;
  for (int k = 1; k < n; ++k) r = r * x;
  return r;
}

__device__ inline double fsign(double a, double b) { return b >= 0.0 ? fabs(a) : -fabs(a); }

__global__ void __launch_bounds__(256) k_tors_terms(const TorsDesc *__restrict__ D) {
#pragma clang fp contract(off)
  const TorsDesc &d = *D;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ntors) return;
  double *tf = d.tf + 12 * (size_t)i;
  double *te = d.te + 2 * (size_t)i;
  for (int k = 0; k < 12; ++k) tf[k] = 0.0;
  te[0] = te[1] = 0.0;
  if (d.tseg) d.tseg[i] = 0.0;
  const double b = d.potb[i];
  if (fabs(b) < 1.0e-12) return;  // fptors.f:92
  const int4 l = d.atoms[i];
  const int type = d.ss[l.x];
  const double x21 = d.x[l.y] - d.x[l.x], y21 = d.y[l.y] - d.y[l.x], z21 = d.z[l.y] - d.z[l.x];
  const double x32 = d.x[l.z] - d.x[l.y], y32 = d.y[l.z] - d.y[l.y], z32 = d.z[l.z] - d.z[l.y];
  const double x43 = d.x[l.w] - d.x[l.z], y43 = d.y[l.w] - d.y[l.z], z43 = d.z[l.w] - d.z[l.z];
  const double rsq21 = x21 * x21 + y21 * y21 + z21 * z21;
  const double rsq32 = x32 * x32 + y32 * y32 + z32 * z32;
  const double rsq43 = x43 * x43 + y43 * y43 + z43 * z43;
  const double rsp21 = sqrt(rsq21), rsp32 = sqrt(rsq32), rsp43 = sqrt(rsq43);
  const double cb1 = (x21 * x32 + y21 * y32 + z21 * z32) / (rsp21 * rsp32);
  const double cb2 = (x43 * x32 + y43 * y32 + z43 * z32) / (rsp43 * rsp32);
  const double cb3 = (x21 * x43 + y21 * y43 + z21 * z43) / (rsp21 * rsp43);
  // fitors.f takes the square roots as they come (no absolute value) and has no singular case
  const double s1 = 1.0 - cb1 * cb1, s2 = 1.0 - cb2 * cb2, s3 = 1.0 - cb3 * cb3;
  const double sb1 = sqrt(d.improper ? s1 : fabs(s1)), sb2 = sqrt(d.improper ? s2 : fabs(s2)),
               sb3 = sqrt(d.improper ? s3 : fabs(s3));
  const double aux = sb1 * sb2;
  if (!d.improper && fabs(aux * sb3) < 1.0e-25) {  // singular angle: the reference perturbs at random (not restated)
    d.status[0] = 1;  // status[0]: singular torsion; status[1]: unsupported multiplicity
    return;
  }
  double coa = (cb1 * cb2 - cb3) / aux;
  const double pa = d.pota[i];
  const int tmp = (int)(pa >= 0.0 ? floor(pa + 0.5) : -floor(-pa + 0.5));  // IDNINT
  const int ntph = abs(tmp);
  const bool harmonic = d.improper && d.potc[i] > 0.0;
  if (ntph == 0 && !harmonic) {
    d.status[1] = 2;
    return;
  }
  const double sig = ntph ? (double)(tmp / ntph) : 0.0;
  const double lim = (double)1.000001f;  // default-real literal (fptors.f:139-140)
  if (coa > 1.0 && coa < lim) coa = 1.0;
  if (coa < -1.0 && coa > -lim) coa = -1.0;
  double aux1 = cb1 * sb2, aux2 = sb1 * cb2;
  const double auxa = -(aux1 * coa + aux2) / aux;
  const double auxb = -(aux2 * coa + aux1) / aux;
  const double auxc = sb3 / aux;
  aux1 = rsp21 * rsp32 * sb1;
  aux2 = rsp32 * rsp43 * sb2;
  const double aux3 = rsp21 * rsp43 * sb3;
  double enrg, qforce;
  const double c2 = ipow_ltr(coa, 2), c3 = ipow_ltr(coa, 3), c4 = ipow_ltr(coa, 4);
  if (harmonic) {  // fitors.f harmonic form on the signed dihedral
    double bb = acos(coa);
    const double xn = y21 * z32 - y32 * z21, yn = z21 * x32 - z32 * x21, zn = x21 * y32 - x32 * y21;
    const double aux4 = fsign(bb, xn * x43 + yn * y43 + zn * z43);
    const double q4 = (double)1.57079f;  // default-real literal
    if (aux4 < 0.0 && fabs(aux4) >= q4) {
      bb = 2.0 * 3.141592653589793 + aux4;
      if (fabs(pa - bb) > 3.14159) bb = -2 * 3.14159 + bb;
    } else if (aux4 < 0.0 && fabs(aux4) < q4) {
      bb = aux4;
      if (fabs(pa - bb) > 3.14159) bb = 2 * 3.14159 + bb;
    } else {
      bb = aux4;
      if (fabs(pa - bb) > 3.14159) bb = 2 * 3.14159 + bb;
      if (fabs(pa - bb) > 3.14159) bb = -4 * 3.14159 + bb;
    }
    if (fabs(pa) < 1.0e-3)
      qforce = 2.0 * b;
    else if (fabs(pa - 3.14159) < 1.0e-3)
      qforce = -2.0 * b;
    else
      qforce = (2.0 * b * (bb - pa)) / sin(bb);
    const double da = bb - pa;
    enrg = b * (da * da);
  } else if (d.improper && (ntph == 5 || ntph > 6)) {  // fitors.f stops
    d.status[1] = 2;
    return;
  } else switch (ntph) {
    case 1: enrg = sig * (fabs(b) - coa * b); qforce = sig * b; break;
    case 2: enrg = sig * (fabs(b) - (2.0 * c2 - 1.0) * b); qforce = sig * (4.0 * b * coa); break;
    case 3: enrg = sig * (fabs(b) - (4.0 * c3 - 3.0 * coa) * b); qforce = sig * ((12.0 * c2 - 3.0) * b); break;
    case 4:
      enrg = sig * (fabs(b) - (8.0 * c4 - 8.0 * c2 + 1.0) * b);
      qforce = sig * ((32.0 * c3 - 16.0 * coa) * b);
      break;
    case 5:
      enrg = sig * (fabs(b) - (16.0 * ipow_ltr(coa, 5) - 20.0 * c3 + 5.0 * coa) * b);
      qforce = sig * ((80.0 * c4 - 60.0 * c2 + 5.0) * b);
      break;
    case 6:
      enrg = sig * (fabs(b) - b * (32.0 * ipow_ltr(coa, 6) - 48.0 * c4 + 18.0 * c2 - 1.0));
      qforce = sig * (b * (192.0 * ipow_ltr(coa, 5) - 192.0 * c3 + 36.0 * coa));
      break;
    case 7:
      enrg = sig * (fabs(b) - b * (64.0 * ipow_ltr(coa, 7) - 112.0 * ipow_ltr(coa, 5) + 56.0 * c3 - 7.0 * coa));
      qforce = sig * (b * (448.0 * ipow_ltr(coa, 6) - 560.0 * c4 + 168.0 * c2 - 7.0));
      break;
    case 8:
      enrg = sig * (fabs(b) - b * (128.0 * ipow_ltr(coa, 8) - 256.0 * ipow_ltr(coa, 6) + 160.0 * c4 - 32.0 * c2 + 1.0));
      qforce = sig * (b * (1024.0 * ipow_ltr(coa, 7) - 1536.0 * ipow_ltr(coa, 5) + 640.0 * c3 - 64.0 * coa));
      break;
    case 12:
    case 14: {
      // fptors.f:251-290: the sine signed by the dihedral's orientation
      const double sin60 = (double)0.866025f;
      const double bb = acos(coa);
      const double xn = y21 * z32 - y32 * z21, yn = z21 * x32 - z32 * x21, zn = x21 * y32 - x32 * y21;
      const double aux4 = fsign(bb, xn * x43 + yn * y43 + zn * z43);
      const double soa = fsign(sqrt(1.0 - coa * coa), aux4);
      // fptors_rem.f:269-273: no derivative term where the sine vanishes
      const double dsoa = (d.rem_variant && fabs(soa) < 1.0e-8) ? 0.0 : -coa / soa;
      double auxn, dauxn;
      if (ntph == 14) {
        enrg = b * (1.0 + 0.5 * (8.0 * c4 - 8.0 * c2 + 1.0));
        auxn = b * sin60 * 4.0 * (soa * c3 - coa * ipow_ltr(soa, 3));
        dauxn = b * sin60 * 4.0 * (dsoa * (c3 - 3.0 * coa * ipow_ltr(soa, 2)) - ipow_ltr(soa, 3) + 3.0 * soa * c2);
        qforce = 0.5 * (32.0 * c3 - 16.0 * coa) * b;
      } else {
        enrg = b * (1.0 + 0.5 * (2.0 * c2 + 1.0));
        auxn = b * sin60 * 2.0 * soa * coa;
        dauxn = b * sin60 * 2.0 * (soa + dsoa * coa);
        qforce = 2.0 * coa * b;
      }
      if (d.rem_variant) {  // fptors_rem.f:278-281, 297-300: no sign factor
        enrg = enrg + auxn;
        qforce = -(qforce + dauxn);
      } else {
        enrg = sig * (enrg + auxn);
        qforce = sig * (-(qforce + dauxn));
      }
      break;
    }
    default: d.status[1] = 2; return;  // the reference stops (fptors.f:291-297)
  }
  te[type - 1] = enrg;
  if (d.remf && d.remf[i]) {  // fptors_rem.f:310-313, fitors_rem.f:206-209, 255-261
    d.tseg[i] = enrg;
    qforce = qforce * d.remf2;
  }
  // fptors.f:152-183 (angle-cosine derivatives) and 299-322 (force)
  const double r21[3] = {x21, y21, z21}, r32[3] = {x32, y32, z32}, r43[3] = {x43, y43, z43};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double d11 = r32[c] / aux1 - cb1 * r21[c] / (rsq21 * sb1);
    const double d13 = -r21[c] / aux1 + cb1 * r32[c] / (rsq32 * sb1);
    const double d12 = -d13 - d11;
    const double d22 = r43[c] / aux2 - cb2 * r32[c] / (rsq32 * sb2);
    const double d24 = -r32[c] / aux2 + cb2 * r43[c] / (rsq43 * sb2);
    const double d23 = -d22 - d24;
    const double d31 = r43[c] / aux3 - cb3 * r21[c] / (rsq21 * sb3);
    const double d33 = r21[c] / aux3 - cb3 * r43[c] / (rsq43 * sb3);
    const double d32 = -d31, d34 = -d33;
    tf[4 * c + 0] = qforce * (auxa * d11 + auxc * d31);
    tf[4 * c + 1] = qforce * (auxa * d12 + auxb * d22 + auxc * d32);
    tf[4 * c + 2] = qforce * (auxa * d13 + auxb * d23 + auxc * d33);
    tf[4 * c + 3] = qforce * (auxb * d24 + auxc * d34);
  }
}

// Bonds (arity 2) and bends (arity 3): the force components of each term in
// the torsion layout (x of l1..l4, then y, z; the unused slots stay zero).
__global__ void __launch_bounds__(256) k_bond_bend_terms(const TorsDesc *__restrict__ D) {
#pragma clang fp contract(off)
  const TorsDesc &d = *D;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.ntors) return;
  double *tf = d.tf + 12 * (size_t)i;
  double *te = d.te + 2 * (size_t)i;
  for (int k = 0; k < 12; ++k) tf[k] = 0.0;
  te[0] = te[1] = 0.0;
  const int4 l = d.atoms[i];
  const int type = d.ss[l.x] - 1;
  if (d.arity == 2) {  // fpbond.f:60-80
    const double x21 = d.x[l.y] - d.x[l.x], y21 = d.y[l.y] - d.y[l.x], z21 = d.z[l.y] - d.z[l.x];
    const double rs21 = sqrt(x21 * x21 + y21 * y21 + z21 * z21);
    const double pa = d.pota[i], pb = d.potb[i];
    const double qforce = -2.0 * pb * (pa - rs21) / rs21;
    const double u[3] = {qforce * x21, qforce * y21, qforce * z21};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      tf[4 * c + 0] = u[c];
      tf[4 * c + 1] = -u[c];
    }
    const double dr = rs21 - pa;
    te[type] = pb * (dr * dr);
    return;
  }
  // fpbend.f:88-215
  const double pb = d.potb[i];
  if (fabs(pb) < 1.0e-12) return;
  const double pa = d.pota[i], pc = d.potc[i], pd = d.potd[i], pe = d.pote[i], pf = d.potf[i], pg = d.potg[i];
  const double r12[3] = {d.x[l.x] - d.x[l.y], d.y[l.x] - d.y[l.y], d.z[l.x] - d.z[l.y]};
  const double r32[3] = {d.x[l.z] - d.x[l.y], d.y[l.z] - d.y[l.y], d.z[l.z] - d.z[l.y]};
  const double rs12 = r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2];
  const double rs32 = r32[0] * r32[0] + r32[1] * r32[1] + r32[2] * r32[2];
  const double dcc2 = sqrt(rs12 * rs32);
  double cb = (r12[0] * r32[0] + r12[1] * r32[1] + r12[2] * r32[2]) / dcc2;
  if (fabs(cb) > 1.0) cb = cb >= 0.0 ? 1.0 : -1.0;
  const double sb = sqrt(1.0 - cb * cb);
  const double bb = acos(cb);
  const double pi = 3.141592653589793;
  const bool linear = fabs(pa - pi) < (double)0.01f;  // default-real literal
  const double qforce = (linear || sb < 1.0e-6) ? -2.0 * pb : 2.0 * pb * (bb - pa) / sb;
  double f1[3], f2[3], f3[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double u1 = r32[c] / dcc2 - cb * r12[c] / rs12;
    const double u3 = r12[c] / dcc2 - cb * r32[c] / rs32;
    f1[c] = qforce * u1;
    f3[c] = qforce * u3;
    f2[c] = qforce * (-u1 - u3);
  }
  double e;
  if (linear) {
    e = 2.0 * pb * (cb + 1.0);
  } else {
    const double da = bb - pa;
    e = pb * (da * da);
  }
  if (fabs(pd) > 1.0e-5 && (fabs(pe) < 1.0e-5 || fabs(pf) > 1.0e-5)) {  // Urey-Bradley 1-3 term
    // fpbend.f:154-156: the 1-3 vector from the coordinates themselves
    const double r31[3] = {d.x[l.z] - d.x[l.x], d.y[l.z] - d.y[l.x], d.z[l.z] - d.z[l.x]};
    const double rsp31 = sqrt(r31[0] * r31[0] + r31[1] * r31[1] + r31[2] * r31[2]);
    const double qf = -2.0 * pd * (pc - rsp31);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      f1[c] += qf * (r31[c] / rsp31);
      f3[c] -= qf * (r31[c] / rsp31);
    }
    const double dr = rsp31 - pc;
    e += pd * (dr * dr);
  }
  bool coupling = false;
  double k12 = 0.0, q1 = 0.0, q2 = 0.0;
  if (fabs(pd) > 1.0e-5 && fabs(pe) > 1.0e-5 && fabs(pf) < 1.0e-5) {
    coupling = true;
    k12 = pd;
    q1 = pc;
    q2 = pe;
  }
  if (fabs(pe) > 1.0e-5 && fabs(pf) > 1.0e-5 && fabs(pg) > 1.0e-5) {
    coupling = true;
    k12 = pe;
    q1 = pf;
    q2 = pg;
  }
  if (coupling) {  // stretch-stretch k12 (r12 - r1)(r32 - r2)
    const double rsp12 = sqrt(rs12), rsp32 = sqrt(rs32);
    const double qf = k12 * (rsp12 - q1), pfc = k12 * (rsp32 - q2);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double a1 = r12[c] / rsp12, a2 = r32[c] / rsp32;
      f1[c] -= pfc * a1;
      f2[c] += pfc * a1 + qf * a2;
      f3[c] -= qf * a2;
    }
    e += k12 * (rsp12 - q1) * (rsp32 - q2);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    tf[4 * c + 0] = f1[c];
    tf[4 * c + 1] = f2[c];
    tf[4 * c + 2] = f3[c];
  }
  te[type] = e;
}

__global__ void __launch_bounds__(256) k_tors_gather(const TorsDesc *__restrict__ D) {
  const TorsDesc &d = *D;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= d.natoms) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int t = d.g_off[a]; t < d.g_off[a + 1]; ++t) {
    const int s = d.g_slot[t], i = s >> 2, k = s & 3;
    sx += d.tf[12 * (size_t)i + k];
    sy += d.tf[12 * (size_t)i + 4 + k];
    sz += d.tf[12 * (size_t)i + 8 + k];
  }
  d.fx[a] = sx;  // every atom's sum is written (zero where no term touches it)
  d.fy[a] = sy;
  d.fz[a] = sz;
}

// d.out[k] = sum over torsions of te[k] (k = 2: of tseg), one block per column, fixed order
__global__ void __launch_bounds__(256) k_tors_energy(const TorsDesc *__restrict__ D) {
  __shared__ double red[4];
  const TorsDesc &d = *D;
  const int k = blockIdx.x;
  double s = 0.0;
  if (k < 2)
    for (int i = threadIdx.x; i < d.ntors; i += blockDim.x) s += d.te[2 * (size_t)i + k];
  else
    for (int i = threadIdx.x; i < d.ntors; i += blockDim.x) s += d.tseg[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) d.out[k] = red[0] + red[1] + red[2] + red[3];
}

void launch_tors(const TorsDesc *d_desc, const TorsDesc &h, hipStream_t st) {
  if (h.ntors > 0) {
    if (h.arity == 4)
      hipLaunchKernelGGL(k_tors_terms, dim3((h.ntors + 255) / 256), dim3(256), 0, st, d_desc);
    else
      hipLaunchKernelGGL(k_bond_bend_terms, dim3((h.ntors + 255) / 256), dim3(256), 0, st, d_desc);
  }
  if (h.ntors > 0 && h.natoms > 0)
    hipLaunchKernelGGL(k_tors_gather, dim3((h.natoms + 255) / 256), dim3(256), 0, st, d_desc);
  hipLaunchKernelGGL(k_tors_energy, dim3(h.tseg ? 3 : 2), dim3(256), 0, st, d_desc);
}

}  // namespace orac
