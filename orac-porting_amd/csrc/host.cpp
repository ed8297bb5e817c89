// host.cpp -- host-side tables and plans of the nonbonded path: FFT plans and
// twiddles, B-spline moduli, linked-cell stencils, exclusion maps, PME descriptors.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include <map>
#include <tuple>

#include "host.hpp"

namespace orac {

thread_local std::string g_last_error;

// Twiddles exp(-2 pi i m / n) in extended precision, rounded once.
static void make_twiddles(int n, std::vector<double2> &t64, std::vector<float2> &t32) {
  t64.resize(n);
  t32.resize(n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    long double a = -two_pi * (long double)m / (long double)n;
    double c = (double)cosl(a), s = (double)sinl(a);
    t64[m] = make_double2(c, s);
    t32[m] = make_float2((float)cosl(a), (float)sinl(a));
  }
}

static bool factor_radices(int n, FftAxis &ax) {
  ax.n = n;
  ax.nr = 0;
  int m = n;
  const int order[4] = {4, 2, 3, 5};
  for (int r : order)
    while (m % r == 0) {
      if (ax.nr >= MAX_RADICES) return false;
      ax.radix[ax.nr++] = r;
      m /= r;
    }
  return m == 1 && n <= MAX_FFT_LEN && n >= 2;
}

// load_bsp_moduli + DFTMOD (pme_fftw3.f90:165-230), with DFTMOD's 2*pi
// built from the default-real (REAL(4)) pi literal.
void bsp_moduli(int order, int nfft, double *out) {
  double arr[MAX_ORDER + 2], darr[MAX_ORDER + 2];
  fill_bspline<double>(0.0, order, arr, darr);
  std::vector<double> b(nfft + order + 2, 0.0);
  for (int i = 2; i <= order + 1; ++i) b[i - 1] = arr[i - 2];
  const double twopi = 2.0 * (double)3.14159265358979323846f;
  for (int k = 1; k <= nfft; ++k) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 1; j <= nfft; ++j) {
      double arg = twopi * (k - 1) * (j - 1) / nfft;
      s1 = s1 + b[j - 1] * std::cos(arg);
      s2 = s2 + b[j - 1] * std::sin(arg);
    }
    out[k - 1] = s1 * s1 + s2 * s2;
  }
  for (int k = 1; k <= nfft; ++k)
    if (out[k - 1] < 1.0e-7) out[k - 1] = 0.5 * (out[k - 2] + out[k]);
}

// dist_ijk.f: corner search and 500-step edge scan between cell (0,0,0) and (ni,nj,nk)
static double dist_ijk(int ni, int nj, int nk, double dx, double dy, double dz, const double *co) {
  static const int nv[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
  const int ndtmax = 500;
  double dmin = 1.0e8, d;
  int imin = 0, jmin = 0;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) {
      double lx = (ni + nv[j][0] - nv[i][0]) * dx, ly = (nj + nv[j][1] - nv[i][1]) * dy,
             lz = (nk + nv[j][2] - nv[i][2]) * dz;
      double mx = co[0] * lx + co[3] * ly + co[6] * lz;
      double my = co[1] * lx + co[4] * ly + co[7] * lz;
      double mz = co[2] * lx + co[5] * ly + co[8] * lz;
      d = mx * mx + my * my + mz * mz;
      if (d < dmin) {
        dmin = d;
        imin = i;
        jmin = j;
      }
    }
  double lx = (ni + nv[jmin][0] - nv[imin][0]) * dx, ly = (nj + nv[jmin][1] - nv[imin][1]) * dy,
         lz = (nk + nv[jmin][2] - nv[imin][2]) * dz;
  double mx = co[0] * lx + co[3] * ly + co[6] * lz;
  double my = co[1] * lx + co[4] * ly + co[7] * lz;
  double mz = co[2] * lx + co[5] * ly + co[8] * lz;
  const double dc[3] = {dx, dy, dz};
  dmin = 1.0e8;
  for (int ax = 0; ax < 3; ++ax) {
    double sgn = (0.5f - (float)nv[jmin][ax]) >= 0.0f ? 1.0 : -1.0;  // sign(1., 0.5-nv), default real
    double dt = sgn * dc[ax] / ndtmax;
    for (int i = 1; i <= ndtmax; ++i) {
      double dmx = i * co[0 + 3 * ax] * dt, dmy = i * co[1 + 3 * ax] * dt, dmz = i * co[2 + 3 * ax] * dt;
      d = (mx + dmx) * (mx + dmx) + (my + dmy) * (my + dmy) + (mz + dmz) * (mz + dmz);
      if (d < dmin) dmin = d;
    }
  }
  return dmin;
}

// lc_index.f:3-125; returns false where the reference STOPs
bool lc_index(int indmax, int ncx, int ncy, int ncz, int32_t *nind, int32_t *indxi, int32_t *indxj,
                     int32_t *indxk, double ctoff, const double *co) {
  double sqcut = ctoff * ctoff;
  double dx = 2.0 / ncx, dy = 2.0 / ncy, dz = 2.0 / ncz;
  int imax = 0, jmax = 0, kmax = 0;
  // Offsets that cannot be in the stencil are not visited (the reference
  // visits every offset; dist_ijk is ~1,600 distance evaluations, so large
  // boxes spent seconds here).  Two points of cells 0 and (i, j, k) differ by
  // at least (|i| - 1) dx in the first reduced coordinate, so their distance
  // is at least (|i| - 1) dx / |row 1 of co^-1|; dist_ijk measures distances
  // between points of the two cells, never less than that.  An offset is
  // skipped only when even |i| - 2 cells of separation reach the cut-off, so
  // the stencil and its order are the reference's.
  int reach[3];
  {
    const double *c = co;  // column-major 3x3
    const double a = c[0], b = c[3], cc = c[6], d = c[1], e = c[4], f = c[7], g = c[2], h = c[5], k = c[8];
    const double det = a * (e * k - f * h) - b * (d * k - f * g) + cc * (d * h - e * g);
    // rows of the inverse, up to the factor 1/det
    const double r[3][3] = {{e * k - f * h, cc * h - b * k, b * f - cc * e},
                            {f * g - d * k, a * k - cc * g, cc * d - a * f},
                            {d * h - e * g, b * g - a * h, a * e - b * d}};
    const double cell[3] = {dx, dy, dz};
    const int nmax[3] = {ncx, ncy, ncz};
    for (int ax = 0; ax < 3; ++ax) {
      const double nrm = std::sqrt(r[ax][0] * r[ax][0] + r[ax][1] * r[ax][1] + r[ax][2] * r[ax][2]) / std::fabs(det);
      const double w = cell[ax] / nrm;  // distance between the planes one cell apart
      reach[ax] = nmax[ax];
      if (std::isfinite(w) && w > 0.0) reach[ax] = std::min(nmax[ax], (int)std::ceil(ctoff / w) + 2);
    }
  }
  nind[0] = 1;
  indxi[0] = indxj[0] = indxk[0] = 0;
  nind[1] = 1;
  indxi[1] = indxj[1] = indxk[1] = 0;
  for (int n = 1; n <= 2; ++n) {
    int istart = n == 1 ? 0 : 1 - ncx;
    for (int i = std::max(istart, -reach[0]); i <= std::min(ncx - 1, reach[0]); ++i) {
      int jstart = (n == 1 && i == 0) ? 0 : 1 - ncy;
      for (int j = std::max(jstart, -reach[1]); j <= std::min(ncy - 1, reach[1]); ++j) {
        int kstart = (n == 1 && i == 0 && j == 0) ? 0 : 1 - ncz;
        for (int k = std::max(kstart, -reach[2]); k <= std::min(ncz - 1, reach[2]); ++k) {
          if (dist_ijk(i, j, k, dx, dy, dz, co) < sqcut) {
            nind[n - 1] += 1;
            if (nind[n - 1] > indmax) return false;
            int m = nind[n - 1];
            indxi[(n - 1) + 2 * (m - 1)] = i;
            indxj[(n - 1) + 2 * (m - 1)] = j;
            indxk[(n - 1) + 2 * (m - 1)] = k;
            imax = std::max(imax, std::abs(i));
            jmax = std::max(jmax, std::abs(j));
            kmax = std::max(kmax, std::abs(k));
            if (i == 0 && j == 0 && k == 0) nind[n - 1] -= 1;
          }
        }
      }
    }
  }
  return !(imax >= (ncx + 1) / 2 || jmax >= (ncy + 1) / 2 || kmax >= (ncz + 1) / 2);
}

double fix_scale_for(const double *q, int n) {
  // every grid point is bounded by sum |q|; keep one spare bit below 2^63
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += std::fabs(q[i]);
  if (s <= 0.0) s = 1.0;
  int e = (int)std::ceil(std::log2(s));
  return std::ldexp(1.0, 61 - e);
}

// symmetric exclusion CSR from the mapnl layout (mappa.f: per atom [m, j > i ...])
void exclusions_from_mapnl(const int32_t *mapnl, int nato, const std::vector<int> &atomg,
                                  std::vector<int> &off, std::vector<int> &idx, std::vector<unsigned char> &inter) {
  std::vector<std::vector<int>> part(nato);
  size_t na = 0;
  for (int i = 0; i < nato; ++i) {
    int m = mapnl[na];
    for (int t = 1; t <= m; ++t) {
      int j = mapnl[na + t] - 1;
      if (j < 0 || j >= nato) throw Error(E_ARG, "mapnl holds an atom index out of range");
      if (j == i) continue;
      part[i].push_back(j);
      part[j].push_back(i);
    }
    na += (size_t)m + 1;
  }
  off.assign(nato + 1, 0);
  inter.assign(nato, 0);
  idx.clear();
  for (int i = 0; i < nato; ++i) {
    auto &v = part[i];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    for (int j : v) {
      idx.push_back(j);
      if (atomg[j] != atomg[i]) inter[i] = 1;
    }
    off[i + 1] = (int)idx.size();
  }
  if (idx.empty()) idx.push_back(0);
}

void build_pme_plan(BufSet &B, const std::string &pfx, PmePlan &P, int k1, int k2, int k3, int order,
                           hipStream_t st) {
  ORAC_REQUIRE(order >= 3 && order <= MAX_ORDER, E_UNSUPPORTED, "spline order must be 3..8");
  int K[3] = {k1, k2, k3};
  for (int d = 0; d < 3; ++d) {
    ORAC_REQUIRE(factor_radices(K[d], P.ax[d]), E_UNSUPPORTED,
                 "grid dimensions must be 2^a 3^b 5^c and at most 256");
    ORAC_REQUIRE(K[d] >= order, E_ARG, "grid dimension smaller than the spline order");
    std::vector<double2> t64;
    std::vector<float2> t32;
    make_twiddles(K[d], t64, t32);
    P.ax[d].tw64 = B.put_vec(pfx + "tw64_" + std::to_string(d), t64, st, false);
    P.ax[d].tw32 = B.put_vec(pfx + "tw32_" + std::to_string(d), t32, st, false);
    P.bsp[d].resize(K[d]);
    bsp_moduli(order, K[d], P.bsp[d].data());
  }
  P.K1 = k1;
  P.K2 = k2;
  P.K3 = k3;
  P.order = order;
}

// mapnl length by walking its per-atom records
size_t mapnl_length(const int32_t *mapnl, int nato) {
  size_t na = 0;
  for (int i = 0; i < nato; ++i) {
    ORAC_REQUIRE(mapnl[na] >= 0, E_ARG, "negative count in mapnl");
    na += (size_t)mapnl[na] + 1;
  }
  return na;
}

CellGrid make_cellgrid(BufSet &B, const std::string &pfx, int ncx, int ncy, int ncz, const int32_t *nind,
                              const int32_t *indxi, const int32_t *indxj, const int32_t *indxk, double rcut,
                              const double *co, int half, hipStream_t st) {
  CellGrid cg;
  cg.ncx = ncx;
  cg.ncy = ncy;
  cg.ncz = ncz;
  cg.nsten = nind[1];
  std::vector<int> sten(3 * (size_t)cg.nsten);
  for (int m = 0; m < cg.nsten; ++m) {
    sten[3 * m] = indxi[1 + 2 * m];
    sten[3 * m + 1] = indxj[1 + 2 * m];
    sten[3 * m + 2] = indxk[1 + 2 * m];
  }
  cg.sten = B.put_vec(pfx + "sten", sten, st, true);
  for (int k = 0; k < 9; ++k) cg.co[k] = co[k];
  cg.sqcut = rcut * rcut;
  cg.half = half;
  return cg;
}

// Device PME descriptor of one system: grids, B-spline moduli, partial-sum
// buffers sized for any tuning, spread bins when the grid admits tiles.
PmeDesc pme_desc_base(BufSet &B, const std::string &pfx, const PmePlan &P, int natoms, const double *oc,
                      double volume, double alpha, double rkcut, bool fp64, const Tuning &tu, hipStream_t st) {
  PmeDesc d;
  std::memset(&d, 0, sizeof d);
  d.natoms = natoms;
  d.K1 = P.K1;
  d.K2 = P.K2;
  d.K3 = P.K3;
  d.order = P.order;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) d.recip[i + 3 * j] = 0.5 * oc[j + 3 * i];
  d.volume = volume;
  d.pi_f32 = (double)3.14159265358979323846f;
  d.fac = d.pi_f32 * d.pi_f32 / (alpha * alpha);
  d.rkcut2 = 0.25 * rkcut * rkcut / (d.pi_f32 * d.pi_f32);
  const int H = P.K1 / 2 + 1;
  const size_t G = (size_t)P.K1 * P.K2 * P.K3, Gc = (size_t)H * P.K2 * P.K3;
  d.Qfix = B.get<unsigned long long>(pfx + "Qfix", G);
  d.FQ = fp64 ? (void *)B.get<double2>(pfx + "FQ", Gc) : (void *)B.get<float2>(pfx + "FQ32", Gc);
  d.Qr = fp64 ? (void *)B.get<double>(pfx + "Qr", G) : (void *)B.get<float>(pfx + "Qr32", G);
  d.bsp1 = B.put_vec(pfx + "bsp1", P.bsp[0], st, true);
  d.bsp2 = B.put_vec(pfx + "bsp2", P.bsp[1], st, true);
  d.bsp3 = B.put_vec(pfx + "bsp3", P.bsp[2], st, true);
  d.ax1 = P.ax[0];
  d.ax2 = P.ax[1];
  d.ax3 = P.ax[2];
  // one K3 pencil per workgroup is the most blocks any tiling uses
  d.conv_part = B.get<double>(pfx + "conv_part", (size_t)H * P.K2 * 8);
  d.gather_part = B.get<double>(pfx + "gather_part", (size_t)std::max(1, (natoms + 31) / 32) * 9);
  // LDS-tiled spread when every axis has a tile edge in [order-1, 8] dividing K;
  // the bins are sized for the smallest tiles (the most bins any tuning uses)
  const int e1 = spread_tile_edge(P.K1, P.order, true), e2 = spread_tile_edge(P.K2, P.order, true),
            e3 = spread_tile_edge(P.K3, P.order, true);
  if (e1 && e2 && e3) {
    const size_t nbin = (size_t)(P.K1 / e1) * (P.K2 / e2) * (P.K3 / e3);
    d.bin_cnt = B.get<int>(pfx + "bin_cnt", nbin);
    d.bin_start = B.get<int>(pfx + "bin_start", nbin);
    d.atom_bin = B.get<int2>(pfx + "atom_bin", std::max(1, natoms));
    d.afr = B.get<double4>(pfx + "afr", std::max(1, natoms));
    d.sfr = B.get<double4>(pfx + "sfr", std::max(1, natoms));
    d.swt = B.get<double>(pfx + "swt", (size_t)std::max(1, natoms) * 3 * P.order);
  }
  pme_apply_tuning(d, tu);
  return d;
}

// Spread path of a tuning: 0 global atomics; 1 tiles of the largest edges
// (the default: every atom's stencil reaches 8 tiles whatever the edge, and
// large tiles give each workgroup the most work); 2 the smallest edges.
void pme_apply_tuning(PmeDesc &d, const Tuning &tu) {
  pme_set_tiles(d, tu.lds_elems);
  d.tiled = 0;
  if (!d.bin_cnt || !tu.spread_tiled || std::getenv("ORAC_HIP_SPREAD_ATOMIC")) return;
  const int K[3] = {d.K1, d.K2, d.K3};
  int big[3], small[3];
  for (int a = 0; a < 3; ++a) {
    big[a] = spread_tile_edge(K[a], d.order, false);
    small[a] = spread_tile_edge(K[a], d.order, true);
  }
  const int *e = tu.spread_tiled == 2 ? small : big;
  d.st1 = e[0];
  d.st2 = e[1];
  d.st3 = e[2];
  d.nb1 = K[0] / e[0];
  d.nb2 = K[1] / e[1];
  d.nb3 = K[2] / e[2];
  d.tiled = 1;
}

double erfc_as(double x) {
  const double a1 = 0.2548296, a2 = -0.28449674, a3 = 1.4214137, a4 = -1.453152, a5 = 1.0614054, qp = 0.3275911;
  double qt = 1.0 / (1.0 + qp * x);
  double expcst = std::exp(-(x * x));
  return ((((a5 * qt + a4) * qt + a3) * qt + a2) * qt + a1) * qt * expcst;
}

// ---------------------------------------------------------------- alchemy

AlchTables alch_tables(int n, const int32_t *ss_alchemy, const double *rmin, const double *epsm, int opls) {
  AlchTables T;
  T.cls.resize(n);
  std::map<std::tuple<double, double, int>, int> ids;
  std::vector<double> cr, ce;
  std::vector<int> cs;
  for (int a = 0; a < n; ++a) {
    const int s = ss_alchemy[a];
    ORAC_REQUIRE(s == 1 || s == 2 || s == -2, E_ARG, "ss_alchemy must be 1, 2 or -2");
    ORAC_REQUIRE(epsm[a] >= 0.0 && std::isfinite(rmin[a]), E_ARG, "bad rmin or epsm");
    auto key = std::make_tuple(rmin[a], epsm[a], s);
    auto it = ids.find(key);
    if (it == ids.end()) {
      it = ids.emplace(key, (int)cr.size()).first;
      cr.push_back(rmin[a]);
      ce.push_back(epsm[a]);
      cs.push_back(s);
    }
    T.cls[a] = it->second;
  }
  T.ncls = (int)cr.size();
  ORAC_REQUIRE(T.ncls <= LJ_DENSE_MAX, E_UNSUPPORTED, "more than 2048 alchemical classes");
  T.pair.resize((size_t)T.ncls * T.ncls);
  for (int i = 0; i < T.ncls; ++i)
    for (int j = 0; j < T.ncls; ++j) {
      // force_alchemy.f:432-437
      const double sigma = opls ? 2.0 * std::pow(cr[j] * cr[i], 0.5) : cr[j] + cr[i];
      const double epsmin = 4.0 * std::sqrt(ce[i] * ce[j]);
      const double s6 = std::pow(sigma, 6.0);
      T.pair[(size_t)i * T.ncls + j] = make_double4(s6 > 0.0 ? 1.0 / s6 : 0.0, epsmin, cs[i] + cs[j] + 4, sigma);
    }
  return T;
}

void alch_set_host(double fact[18], double ladd, double lrm, bool intramol, bool ljintra, bool qq) {
  for (int k = 0; k < 18; ++k) fact[k] = 0.0;
  auto F = [&](int t, int k) -> double & { return fact[2 * (t + 4) + k]; };
  const double a = std::fabs(ladd), r = std::fabs(lrm);
  F(2, 0) = 0.0; F(2, 1) = 1.0;    // normal-normal
  F(3, 0) = a; F(3, 1) = 1.0 - a;  // normal-added
  F(-1, 0) = r; F(-1, 1) = 1.0 - r;  // normal-removed
  if (intramol && qq) {
    F(4, 0) = a * a; F(4, 1) = 1.0 - a * a;
    F(-4, 0) = r * r; F(-4, 1) = 1.0 - r * r;
  } else if (intramol && !ljintra) {
    F(4, 0) = a; F(4, 1) = 1.0 - a;
    F(-4, 0) = r; F(-4, 1) = 1.0 - r;
  } else {
    F(4, 0) = 0.0; F(4, 1) = 1.0;
    F(-4, 0) = 0.0; F(-4, 1) = 1.0;
  }
  F(0, 0) = 0.0; F(0, 1) = 0.0;    // added-removed
}

void alch_factors(const double lam[8], bool intramol, bool ljintra, double qx, double ljx, bool shifted,
                  double out[72]) {
  double fl[18], fl0[18], fq[18], fq0[18];
  alch_set_host(fl, lam[0], lam[1], intramol, ljintra, false);
  alch_set_host(fl0, lam[2], lam[3], intramol, ljintra, false);
  alch_set_host(fq, lam[4], lam[5], intramol, ljintra, true);
  alch_set_host(fq0, lam[6], lam[7], intramol, ljintra, true);
  // two6 = 2**(1./6.) with the REAL(4) exponent (force_alchemy_shifted.f:178)
  const double lj = shifted ? ljx * std::pow(2.0, (double)(1.0f / 6.0f)) : ljx;
  for (int t = 0; t < 9; ++t) {
    double *o = out + 8 * t;
    o[0] = lj * fl[2 * t];
    o[1] = fl[2 * t + 1];
    o[2] = lj * fl0[2 * t];
    o[3] = fl0[2 * t + 1];
    o[4] = shifted ? qx * fq[2 * t] : qx * (fq[2 * t] * fq[2 * t]);
    o[5] = fq[2 * t + 1];
    o[6] = shifted ? qx * fq0[2 * t] : qx * (fq0[2 * t] * fq0[2 * t]);
    o[7] = fq0[2 * t + 1];
  }
}

void alch_charges(int n, const double *q, const int32_t *ss_alchemy, double laddq, double lrmq, double *out) {
  for (int a = 0; a < n; ++a) {
    const double l = ss_alchemy[a] == 2 ? laddq : (ss_alchemy[a] == -2 ? -lrmq : 0.0);
    const double m = std::fabs(l);
    out[a] = m == 1.0 ? 0.0 : (m == 0.0 ? q[a] : (1. - m) * q[a]);
  }
}

// ---------------------------------------------------------------- ERF_CORR tables

// Cubic spline interpolation coefficients, C. de Boor, "A Practical Guide to
// Splines", routine CUBSPL (the form erfc_spline.f:71-193 carries).  c is
// (4, n) column-major with the values in c(1, i) on entry; on exit c(2..4, i)
// hold the first derivative, the second derivative and the third derivative
// on [tau(i), tau(i+1)].  ibcbeg/ibcend: 0 not-a-knot, 1 slope given in
// c(2, 1)/c(2, n), 2 second derivative given there.
void cubspl(const double *tau, double *c, int n, int ibcbeg, int ibcend) {
  auto C = [&](int k, int i) -> double & { return c[(k - 1) + 4 * (i - 1)]; };
  auto T = [&](int i) { return tau[i - 1]; };
  const int l = n - 1;
  for (int m = 2; m <= n; ++m) {
    C(3, m) = T(m) - T(m - 1);
    C(4, m) = (C(1, m) - C(1, m - 1)) / C(3, m);
  }
  double g = 0.0;
  bool skip_interior = false;  // n == 2: straight to the end condition
  if (ibcbeg == 0) {
    if (n > 2) {
      C(4, 1) = C(3, 3);
      C(3, 1) = C(3, 2) + C(3, 3);
      C(2, 1) = ((C(3, 2) + 2.0 * C(3, 1)) * C(4, 2) * C(3, 3) + C(3, 2) * C(3, 2) * C(4, 3)) / C(3, 1);
    } else {
      C(4, 1) = 1.0;
      C(3, 1) = 1.0;
      C(2, 1) = 2.0 * C(4, 2);
      skip_interior = true;
    }
  } else if (ibcbeg == 1) {
    C(4, 1) = 1.0;
    C(3, 1) = 0.0;
    if (n == 2) skip_interior = true;
  } else {
    C(4, 1) = 2.0;
    C(3, 1) = 1.0;
    C(2, 1) = 3.0 * C(4, 2) - C(3, 2) / 2.0 * C(2, 1);
    if (n == 2) skip_interior = true;
  }
  // 0: back substitution only; 1: g = -1 / c(4, n-1), then the last row; 2: the last row with g set
  int last = 0;
  if (!skip_interior) {
    for (int m = 2; m <= l; ++m) {
      g = -C(3, m + 1) / C(4, m - 1);
      C(2, m) = g * C(2, m - 1) + 3.0 * (C(3, m) * C(4, m + 1) + C(3, m + 1) * C(4, m));
      C(4, m) = g * C(3, m - 1) + 2.0 * (C(3, m) + C(3, m + 1));
    }
    if (ibcend == 0) {
      if (n == 3 && ibcbeg == 0) {
        C(2, n) = 2.0 * C(4, n);
        C(4, n) = 1.0;
        last = 1;
      } else {
        g = C(3, n - 1) + C(3, n);
        C(2, n) = ((C(3, n) + 2.0 * g) * C(4, n) * C(3, n - 1) +
                   C(3, n) * C(3, n) * (C(1, n - 1) - C(1, n - 2)) / C(3, n - 1)) / g;
        g = -g / C(4, n - 1);
        C(4, n) = C(3, n - 1);
        last = 2;
      }
    } else if (ibcend == 2) {
      C(2, n) = 3.0 * C(4, n) + C(3, n) / 2.0 * C(2, n);
      C(4, n) = 2.0;
      last = 1;
    }
  } else {
    if (ibcend == 0) {
      if (ibcbeg > 0) {
        C(2, n) = 2.0 * C(4, n);
        C(4, n) = 1.0;
        last = 1;
      } else {
        C(2, n) = C(4, n);
      }
    } else if (ibcend == 2) {
      C(2, n) = 3.0 * C(4, n) + C(3, n) / 2.0 * C(2, n);
      C(4, n) = 2.0;
      last = 1;
    }
  }
  if (last == 1) g = -1.0 / C(4, n - 1);
  if (last != 0) {
    C(4, n) = g * C(3, n - 1) + C(4, n);
    C(2, n) = (g * C(2, n - 1) + C(2, n)) / C(4, n);
  }
  for (int j = l; j > 0; --j) C(2, j) = (C(2, j) - C(3, j) * C(2, j + 1)) / C(4, j);
  for (int i = 2; i <= n; ++i) {
    const double dtau = C(3, i);
    const double divdf1 = (C(1, i) - C(1, i - 1)) / dtau;
    const double divdf3 = C(2, i - 1) + C(2, i) - 2.0 * divdf1;
    C(3, i - 1) = 2.0 * (divdf1 - C(2, i - 1) - divdf3) / dtau;
    C(4, i - 1) = (divdf3 / dtau) * (6.0 / dtau);
  }
}

// Integrand of the reciprocal-cutoff correction (int_corr_erf_spline.f:70-80):
// exp(-k^2 / 4 alpha^2) sin(k r) / (k r)
static double erf_corr_func(double x, double alpha, double r) {
  if (r < 1.0e-8) return std::exp(-(x * x) / (4.0 * (alpha * alpha)));
  return std::exp(-(x * x) / (4.0 * (alpha * alpha))) * std::sin(r * x) / (r * x);
}

// Simpson's rule by refinement of the extended trapezoid rule (the published
// routines qsimp/trapzd, Numerical Recipes section 4.2), as
// int_corr_erf_spline.f:37-69 uses them: tolerance 1e-6, at most 20 levels.
static double qsimp_erf(double a, double b, double alpha, double r) {
  double st = 0.0, ost = -1.0e30, os = -1.0e30, s = 0.0;
  long it = 1;
  for (int j = 1; j <= 20; ++j) {
    if (j == 1) {
      st = 0.5 * (b - a) * (erf_corr_func(a, alpha, r) + erf_corr_func(b, alpha, r));
      it = 1;
    } else {
      const double tnm = (double)it;
      const double del = (b - a) / tnm;
      double x = a + 0.5 * del, sum = 0.0;
      for (long k = 1; k <= it; ++k) {
        sum = sum + erf_corr_func(x, alpha, r);
        x = x + del;
      }
      st = 0.5 * (st + (b - a) * sum / tnm);
      it = 2 * it;
    }
    s = (4.0 * st - ost) / 3.0;
    if (std::fabs(s - os) < 1.0e-6 * std::fabs(os)) return s;
    os = s;
    ost = st;
  }
  return s;
}

void int_corr_erf_spline(double rmin, double rmax, int nbin, double alpha, double kcut, double *erf_arr_corr,
                         double *tau) {
  const double del = (rmax - rmin) / (double)(float)nbin;
  const double pi = 3.14159265358979323846;
  const double kcut_max = 10.0 * 2.0 * alpha;
  for (int i = 1; i <= nbin + 1; ++i) {
    const double r = del * (i - 1) + rmin;
    const double s = qsimp_erf(kcut, kcut_max, alpha, r);
    tau[i - 1] = r;
    erf_arr_corr[4 * (i - 1)] = 2.0 * s / pi;
  }
  cubspl(tau, erf_arr_corr, nbin + 1, 0, 0);
}

double erf_corr_cutoff(const double *oc, int nfft1, int nfft2, int nfft3) {
  const double pi = 4.0 * std::atan(1.0);
  const double aux = 0.5 * 0.5 * 2.0 * pi;
  const double m1 = aux * (oc[0] * nfft1 + oc[1] * nfft2 + oc[2] * nfft3);
  double rkcut = m1;
  const double m2 = aux * (oc[3] * nfft1 + oc[4] * nfft2 + oc[5] * nfft3);
  if (m2 < m1) rkcut = m2;
  const double m3 = aux * (oc[6] * nfft1 + oc[7] * nfft2 + oc[8] * nfft3);
  if (m3 < m2) rkcut = m3;
  return (double)0.9f * rkcut;  // default-real literal (erf_corr_cutoff.f:49)
}

}  // namespace orac
