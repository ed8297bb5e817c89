// api.cpp -- host runtime and C ABI (include/orac_hip.h) of the gfx950 nonbonded path.
//
// A context owns one device, one HIP stream and grow-only device buffers.
// Plain-mode entry points convert the reference's arrays to the device
// layouts, run the kernels and copy results back with the reference's
// accumulate/overwrite semantics.  Static topology arrays are compared with
// the copy last uploaded and sent again only when their bytes change.
// Resident systems keep topology, coordinates, lists and grids in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/orac_hip.h"
#include "common.hpp"
#include "kernels.hpp"

namespace orac {

static thread_local std::string g_last_error;

// ---------------------------------------------------------------- device buffers

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void *ensure(size_t n) {
    if (n == 0) n = 8;
    if (n > bytes) {
      if (p) ORAC_HIP(hipFree(p));
      p = nullptr;
      size_t cap = std::max(n, bytes + bytes / 4);
      ORAC_HIP(hipMalloc(&p, cap));
      bytes = cap;
    }
    return p;
  }
};

struct BufSet {
  std::map<std::string, std::unique_ptr<DevBuf>> bufs;
  std::map<std::string, std::vector<char>> cache;
  template <class T>
  T *get(const std::string &name, size_t n) {
    auto &b = bufs[name];
    if (!b) b.reset(new DevBuf());
    return static_cast<T *>(b->ensure(n * sizeof(T)));
  }
  // upload n elements; cached uploads are skipped when the bytes are unchanged
  template <class T>
  T *put(const std::string &name, const T *host, size_t n, hipStream_t st, bool cached) {
    T *d = get<T>(name, n);
    size_t nb = n * sizeof(T);
    if (cached) {
      auto &c = cache[name];
      if (c.size() == nb && nb && std::memcmp(c.data(), host, nb) == 0) return d;
      c.assign(reinterpret_cast<const char *>(host), reinterpret_cast<const char *>(host) + nb);
    }
    if (nb) ORAC_HIP(hipMemcpyAsync(d, host, nb, hipMemcpyHostToDevice, st));
    return d;
  }
  template <class T>
  T *put_vec(const std::string &name, const std::vector<T> &v, hipStream_t st, bool cached) {
    return put<T>(name, v.data(), v.size(), st, cached);
  }
};

// ---------------------------------------------------------------- host-side tables

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
static void bsp_moduli(int order, int nfft, double *out) {
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
static bool lc_index(int indmax, int ncx, int ncy, int ncz, int32_t *nind, int32_t *indxi, int32_t *indxj,
                     int32_t *indxk, double ctoff, const double *co) {
  double sqcut = ctoff * ctoff;
  double dx = 2.0 / ncx, dy = 2.0 / ncy, dz = 2.0 / ncz;
  int imax = 0, jmax = 0, kmax = 0;
  nind[0] = 1;
  indxi[0] = indxj[0] = indxk[0] = 0;
  nind[1] = 1;
  indxi[1] = indxj[1] = indxk[1] = 0;
  for (int n = 1; n <= 2; ++n) {
    int istart = n == 1 ? 0 : 1 - ncx;
    for (int i = istart; i <= ncx - 1; ++i) {
      int jstart = (n == 1 && i == 0) ? 0 : 1 - ncy;
      for (int j = jstart; j <= ncy - 1; ++j) {
        int kstart = (n == 1 && i == 0 && j == 0) ? 0 : 1 - ncz;
        for (int k = kstart; k <= ncz - 1; ++k) {
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

static double fix_scale_for(const double *q, int n) {
  // every grid point is bounded by sum |q|; keep one spare bit below 2^63
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += std::fabs(q[i]);
  if (s <= 0.0) s = 1.0;
  int e = (int)std::ceil(std::log2(s));
  return std::ldexp(1.0, 61 - e);
}

// symmetric exclusion CSR from the mapnl layout (mappa.f: per atom [m, j > i ...])
static void exclusions_from_mapnl(const int32_t *mapnl, int nato, const std::vector<int> &atomg,
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

// ---------------------------------------------------------------- PME plan

struct PmePlan {
  int K1 = 0, K2 = 0, K3 = 0, order = 0;
  FftAxis ax[3];
  std::vector<double> bsp[3];
};

static void build_pme_plan(BufSet &B, const std::string &pfx, PmePlan &P, int k1, int k2, int k3, int order,
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

static PmeDesc pme_desc_base(BufSet &B, const std::string &pfx, const PmePlan &P, int natoms, const double *oc,
                             double volume, double alpha, double rkcut, bool fp64, int lds_elems, hipStream_t st) {
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
  const size_t G = (size_t)P.K1 * P.K2 * P.K3, Gc = (size_t)(P.K1 / 2 + 1) * P.K2 * P.K3;
  d.Qfix = B.get<unsigned long long>(pfx + "Qfix", G);
  d.FQ = fp64 ? (void *)B.get<double2>(pfx + "FQ", Gc) : (void *)B.get<float2>(pfx + "FQ32", Gc);
  d.Qr = fp64 ? (void *)B.get<double>(pfx + "Qr", G) : (void *)B.get<float>(pfx + "Qr32", G);
  d.bsp1 = B.put_vec(pfx + "bsp1", P.bsp[0], st, true);
  d.bsp2 = B.put_vec(pfx + "bsp2", P.bsp[1], st, true);
  d.bsp3 = B.put_vec(pfx + "bsp3", P.bsp[2], st, true);
  d.ax1 = P.ax[0];
  d.ax2 = P.ax[1];
  d.ax3 = P.ax[2];
  d.conv_nblk = conv_blocks(P.K1, P.K2, P.K3, lds_elems);
  d.conv_part = B.get<double>(pfx + "conv_part", (size_t)d.conv_nblk * 8);
  d.gather_nblk = (natoms + 31) / 32;  // k_gather: 8 lanes per atom, 32 atoms per block
  d.gather_part = B.get<double>(pfx + "gather_part", (size_t)std::max(1, d.gather_nblk) * 9);
  // LDS-tiled spread when every axis has a tile edge in [order-1, 8] dividing K
  d.st1 = spread_tile_edge(P.K1, P.order);
  d.st2 = spread_tile_edge(P.K2, P.order);
  d.st3 = spread_tile_edge(P.K3, P.order);
  if (!d.st1 || !d.st2 || !d.st3 || std::getenv("ORAC_HIP_SPREAD_ATOMIC")) d.st1 = d.st2 = d.st3 = 0;
  if (d.st1) {
    d.nb1 = P.K1 / d.st1;
    d.nb2 = P.K2 / d.st2;
    d.nb3 = P.K3 / d.st3;
    const size_t nbin = (size_t)d.nb1 * d.nb2 * d.nb3;
    d.bin_cnt = B.get<int>(pfx + "bin_cnt", nbin);
    d.bin_start = B.get<int>(pfx + "bin_start", nbin);
    d.atom_bin = B.get<int2>(pfx + "atom_bin", std::max(1, natoms));
    d.afr = B.get<double4>(pfx + "afr", std::max(1, natoms));
    d.sfr = B.get<double4>(pfx + "sfr", std::max(1, natoms));
  }
  return d;
}

// ---------------------------------------------------------------- resident system

enum Stage { ST_LIST, ST_CENTRES, ST_PAIRS, ST_PME, ST_CORR, ST_END, NSTAGE };

struct System {
  BufSet B;
  std::string pfx = "s_";
  int natoms = 0, ngrp = 0, nmol = 0, ntypes = 0, max_group = 1;
  int precision = 0;
  double co[9], oc[9], volume = 0, alpha = 0;
  double rcut = 0, rtol = 0, rnei = 0;
  PmePlan plan;
  PmeDesc pme_h;
  PmeDesc *pme_d = nullptr;
  PairDesc pair_h;
  PairDesc *pair_d = nullptr;
  CorrDesc corr_h;
  CorrDesc *corr_d = nullptr;
  CellGrid cg;
  int ncell = 0;
  size_t list_cap = 0;
  int *list_total = nullptr;   // device: [0] total entries, [1] overflow flag
  double *xp = nullptr, *yp = nullptr, *zp = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;
  double *out = nullptr;       // device: pme[20], pairs[16], corr[4], self[2]
  std::vector<int> mfirst, mlast;
  double self_e[2] = {0, 0};
  bool have_list = false;
  bool profile = false;
  hipEvent_t ev[NSTAGE] = {};
  int32_t last_pairs = 0;
};

}  // namespace orac

struct orac_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  orac::BufSet B;
  orac::PmePlan plan;
  bool plan_ok = false;
  std::map<int, std::unique_ptr<orac::System>> systems;
  int next_sys = 1;
  orac::PmeTiles tiles;
};

namespace orac {

template <class F>
static int guarded(F &&f) {
  try {
    f();
    return OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return E_HIP;
  } catch (...) {
    g_last_error = "unknown error";
    return E_HIP;
  }
}

static void set_device(orac_hip_ctx *c) {
  ORAC_REQUIRE(c != nullptr, E_ARG, "null context");
  ORAC_HIP(hipSetDevice(c->device));
}

// mapnl length by walking its per-atom records
static size_t mapnl_length(const int32_t *mapnl, int nato) {
  size_t na = 0;
  for (int i = 0; i < nato; ++i) {
    ORAC_REQUIRE(mapnl[na] >= 0, E_ARG, "negative count in mapnl");
    na += (size_t)mapnl[na] + 1;
  }
  return na;
}

}  // namespace orac

using namespace orac;

// ================================================================ C ABI

extern "C" int orac_hip_last_error(char *buf, int32_t len) {
  if (!buf || len <= 0) return E_ARG;
  std::strncpy(buf, g_last_error.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return OK;
}

extern "C" int orac_hip_create(orac_hip_ctx **ctx, int32_t device) {
  return guarded([&] {
    ORAC_REQUIRE(ctx != nullptr, E_ARG, "null output pointer");
    int n = 0;
    ORAC_HIP(hipGetDeviceCount(&n));
    ORAC_REQUIRE(device >= 0 && device < n, E_ARG, "no such HIP device");
    auto c = std::make_unique<orac_hip_ctx>();
    c->device = device;
    ORAC_HIP(hipSetDevice(device));
    ORAC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *ctx = c.release();
  });
}

extern "C" int orac_hip_destroy(orac_hip_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->systems)
      for (auto &e : kv.second->ev)
        if (e) (void)hipEventDestroy(e);
    ctx->systems.clear();
    ctx->B.bufs.clear();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
  });
}

extern "C" int orac_hip_invalidate(orac_hip_ctx *ctx) {
  return guarded([&] {
    ORAC_REQUIRE(ctx, E_ARG, "null context");
    ctx->B.cache.clear();
  });
}

extern "C" int orac_hip_device_name(orac_hip_ctx *ctx, char *buf, int32_t len) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(buf && len > 0, E_ARG, "bad buffer");
    hipDeviceProp_t p;
    ORAC_HIP(hipGetDeviceProperties(&p, ctx->device));
    std::strncpy(buf, p.gcnArchName, (size_t)len - 1);
    buf[len - 1] = 0;
  });
}

extern "C" int orac_hip_sync(orac_hip_ctx *ctx) {
  return guarded([&] {
    set_device(ctx);
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int orac_hip_stream(orac_hip_ctx *ctx, void **stream) {
  return guarded([&] {
    ORAC_REQUIRE(ctx && stream, E_ARG, "null argument");
    *stream = (void *)ctx->stream;
  });
}

extern "C" int orac_hip_decide_force(double rinn, double alphal, int32_t *lskip_ewald) {
  return guarded([&] {
    ORAC_REQUIRE(lskip_ewald, E_ARG, "null output");
    const double a1 = 0.2548296, a2 = -0.28449674, a3 = 1.4214137, a4 = -1.453152, a5 = 1.0614054, qp = 0.3275911;
    double qt = 1.0 / (1.0 + qp * alphal * rinn);
    double expcst = std::exp(-alphal * rinn * alphal * rinn);
    double erfcst = ((((a5 * qt + a4) * qt + a3) * qt + a2) * qt + a1) * qt * expcst;
    *lskip_ewald = erfcst < 10.0e-4 ? 1 : 0;
  });
}

// ---------------------------------------------------------------- PME

extern "C" int orac_hip_fft_pme_init(orac_hip_ctx *ctx, int32_t numatoms, int32_t nfft1, int32_t nfft2,
                                     int32_t nfft3, int32_t order, double *bsp_mod1, double *bsp_mod2,
                                     double *bsp_mod3) {
  return guarded([&] {
    set_device(ctx);
    (void)numatoms;
    build_pme_plan(ctx->B, "p_", ctx->plan, nfft1, nfft2, nfft3, order, ctx->stream);
    ctx->plan_ok = true;
    if (bsp_mod1) std::copy(ctx->plan.bsp[0].begin(), ctx->plan.bsp[0].end(), bsp_mod1);
    if (bsp_mod2) std::copy(ctx->plan.bsp[1].begin(), ctx->plan.bsp[1].end(), bsp_mod2);
    if (bsp_mod3) std::copy(ctx->plan.bsp[2].begin(), ctx->plan.bsp[2].end(), bsp_mod3);
  });
}

extern "C" int orac_hip_fft_pme(orac_hip_ctx *ctx, int32_t ntap, const double *xp, const double *yp,
                                const double *zp, const double *xpcm, const double *ypcm, const double *zpcm,
                                const double *pmechg, const double *co, const double *oc, double volume,
                                double alphal, int32_t order, int32_t nfft1, int32_t nfft2, int32_t nfft3,
                                double *eer, double *fpx, double *fpy, double *fpz, double *stressc,
                                const int32_t *atomp, const int32_t *grppt, int32_t pressure, double rkcut) {
  return guarded([&] {
    set_device(ctx);
    (void)grppt;
    ORAC_REQUIRE(ntap >= 0, E_ARG, "negative atom count");
    ORAC_REQUIRE(ctx->plan_ok && ctx->plan.K1 == nfft1 && ctx->plan.K2 == nfft2 && ctx->plan.K3 == nfft3 &&
                     ctx->plan.order == order,
                 E_STATE, "fft_pme called with a grid or order different from fft_pme_init");
    ORAC_REQUIRE(eer && fpx && fpy && fpz && stressc && xp && yp && zp && pmechg && co && oc, E_ARG,
                 "null array argument");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    PmeDesc d = pme_desc_base(B, "p_", ctx->plan, ntap, oc, volume, alphal, rkcut, true, ctx->tiles.lds_elems, st);
    d.xp = B.put("p_xp", xp, ntap, st, false);
    d.yp = B.put("p_yp", yp, ntap, st, false);
    d.zp = B.put("p_zp", zp, ntap, st, false);
    d.co = B.put("p_co", co, 9, st, false);
    d.q = B.put("p_q", pmechg, ntap, st, true);
    d.fixscale = fix_scale_for(pmechg, ntap);
    d.fixinv = 1.0 / d.fixscale;
    d.fx = B.get<double>("p_fx", ntap);
    d.fy = B.get<double>("p_fy", ntap);
    d.fz = B.get<double>("p_fz", ntap);
    std::vector<int> ap;
    if (pressure) {
      ORAC_REQUIRE(atomp && xpcm && ypcm && zpcm, E_ARG, "pressure needs atomp and xpcm");
      ap.resize(ntap);
      int nprot = 0;
      for (int i = 0; i < ntap; ++i) {
        ap[i] = atomp[i] - 1;
        ORAC_REQUIRE(ap[i] >= 0, E_ARG, "atomp holds an index below 1");
        nprot = std::max(nprot, atomp[i]);
      }
      d.atomp = B.put_vec("p_atomp", ap, st, true);
      d.xpcm = B.put("p_xpcm", xpcm, nprot, st, false);
      d.ypcm = B.put("p_ypcm", ypcm, nprot, st, false);
      d.zpcm = B.put("p_zpcm", zpcm, nprot, st, false);
    }
    PmeDesc *dd = B.put("p_desc", &d, 1, st, false);
    double *dout = B.get<double>("p_out", 20);
    launch_pme<double>(dd, d, 1, dout, st, ctx->tiles);
    ORAC_HIP(hipGetLastError());
    std::vector<double> fx(ntap), fy(ntap), fz(ntap);
    double out[20];
    ORAC_HIP(hipMemcpyAsync(fx.data(), d.fx, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fy.data(), d.fy, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fz.data(), d.fz, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(out, dout, sizeof out, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    *eer = out[0];
    if (pressure) {
      // stressc(r,c) = stressc(r,c) - vir(r,c) + stc(3r+c)   (fft_pme.f:97-105)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) stressc[r + 3 * c] = stressc[r + 3 * c] - out[1 + r + 3 * c] + out[10 + 3 * r + c];
    }
    for (int i = 0; i < ntap; ++i) {
      fpx[i] = fpx[i] + fx[i];
      fpy[i] = fpy[i] + fy[i];
      fpz[i] = fpz[i] + fz[i];
    }
  });
}

// ---------------------------------------------------------------- lists

extern "C" int orac_hip_lc_index(int32_t indmax, int32_t ncx, int32_t ncy, int32_t ncz, int32_t *nind,
                                 int32_t *indxi, int32_t *indxj, int32_t *indxk, double ctoff, const double *co) {
  return guarded([&] {
    ORAC_REQUIRE(nind && indxi && indxj && indxk && co, E_ARG, "null array argument");
    ORAC_REQUIRE(ncx > 0 && ncy > 0 && ncz > 0 && indmax > 0, E_ARG, "cell counts must be positive");
    ORAC_REQUIRE(lc_index(indmax, ncx, ncy, ncz, nind, indxi, indxj, indxk, ctoff, co), E_CAPACITY,
                 "lc_index: stencil exceeds indmax or a cell would be visited twice (decrease the cutoff or "
                 "increase the cell counts)");
  });
}

static CellGrid make_cellgrid(BufSet &B, const std::string &pfx, int ncx, int ncy, int ncz, const int32_t *nind,
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

extern "C" int orac_hip_lc_list(orac_hip_ctx *ctx, int32_t ncx, int32_t ncy, int32_t ncz, const int32_t *nind,
                                const int32_t *indxi, const int32_t *indxj, const int32_t *indxk, double rcut,
                                const double *co, const double *xp, const double *yp, const double *zp,
                                int32_t natp, int32_t *nnlpp, int64_t nnlpp_capacity, int64_t *nentries) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(nind && indxi && indxj && indxk && co && xp && yp && zp && nnlpp, E_ARG, "null array argument");
    ORAC_REQUIRE(ncx > 0 && ncy > 0 && ncz > 0 && natp >= 0 && nind[1] >= 1, E_ARG, "bad cell grid");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    CellGrid cg = make_cellgrid(B, "l_", ncx, ncy, ncz, nind, indxi, indxj, indxk, rcut, co, 1, st);
    cg.cpos = B.get<double4>("l_cpos", natp);
    int ncell = ncx * ncy * ncz;
    double *dx = B.put("l_x", xp, natp, st, false), *dy = B.put("l_y", yp, natp, st, false),
           *dz = B.put("l_z", zp, natp, st, false);
    int *cell_of = B.get<int>("l_cellof", natp), *cnt = B.get<int>("l_cnt", ncell),
        *start = B.get<int>("l_start", ncell), *cur = B.get<int>("l_cur", ncell),
        *items = B.get<int>("l_items", natp);
    launch_cell_build(natp, dx, dy, dz, cg, cell_of, cnt, start, cur, items, st);
    int *count = B.get<int>("l_count", natp + 1), *off = B.get<int>("l_off", natp + 1),
        *total = B.get<int>("l_total", 2);
    launch_list_count(natp, dx, dy, dz, cg, cell_of, start, cnt, items, count, st);
    launch_exclusive_scan(count, off, natp, total, st);
    int tot = 0;
    ORAC_HIP(hipMemcpyAsync(&tot, total, sizeof(int), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    int64_t need = (int64_t)tot + natp;
    if (nentries) *nentries = need;
    ORAC_REQUIRE(need <= nnlpp_capacity, E_CAPACITY,
                 "lc_list: dimensions of NNLPP are insufficient (" + std::to_string(need) + " entries needed)");
    int *out = B.get<int>("l_out", (size_t)need);
    launch_list_write(natp, dx, dy, dz, cg, cell_of, start, cnt, items, off, out, 1, st);
    ORAC_HIP(hipGetLastError());
    ORAC_HIP(hipMemcpyAsync(nnlpp, out, (size_t)need * sizeof(int), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
  });
}

// ---------------------------------------------------------------- force

extern "C" int orac_hip_force(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                              const double *zp0, const double *xpg, const double *ypg, const double *zpg,
                              const double *charge, const int32_t *nbtype, const int32_t *type, int32_t ma,
                              int32_t nato, const int32_t *atomg, const double *xpcm, const double *ypcm,
                              const double *zpcm, const int32_t *groupp, const int32_t *atomp, const double *co,
                              const double *ecc12, const double *ecc6, double alphal, const int32_t *mapnl,
                              int32_t ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv,
                              double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx,
                              double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1,
                              int64_t nnlpp1_capacity, double rneigh, double rinn, double rout, double *rtolinn,
                              double rtolout, const double *gmass, int32_t iz, int32_t *ncount1, const double *mass,
                              double massinfty, int32_t lskip_ewald, int32_t ewald) {
  return guarded([&] {
    set_device(ctx);
    (void)groupp;
    (void)atomg;
    ORAC_REQUIRE(nato >= 0 && ngrp >= 0 && ma >= 1, E_ARG, "bad sizes");
    ORAC_REQUIRE(ss_index && xp0 && yp0 && zp0 && xpg && ypg && zpg && charge && nbtype && type && co && ecc12 &&
                     ecc6 && mapnl && grppt && rtolinn && ncount1 && nnlpp && gmass && mass,
                 E_ARG, "null array argument");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    // force.f:145-166
    const double r2inn = rinn * rinn, r2out = rout * rout;
    const double rinn0 = rinn + *rtolinn, r2inn0 = rinn0 * rinn0;
    if (rinn < 1.0e-5) *rtolinn = 0.0;
    const double rout0 = rout + rtolout, r2out0 = rout0 * rout0;
    const double r2neigh = (rinn0 + rneigh) * (rinn0 + rneigh);
    *ncount1 = 0;
    const bool compute = !(iz == 0 || rinn0 >= rout);

    // groups
    for (int a = 0; a < nato; ++a)
      ORAC_REQUIRE(ss_index[a] == 1 || ss_index[a] == 2, E_ARG, "ss_index must be 1 (solute) or 2 (solvent)");
    std::vector<int> gfirst(ngrp), glast(ngrp), gss(ngrp), agroup(nato, -1);
    int max_group = 1;
    for (int g = 0; g < ngrp; ++g) {
      gfirst[g] = grppt[2 * g] - 1;
      glast[g] = grppt[2 * g + 1] - 1;
      ORAC_REQUIRE(gfirst[g] >= 0 && glast[g] >= gfirst[g] && glast[g] < nato, E_ARG, "bad grppt");
      gss[g] = ss_index[gfirst[g]];
      max_group = std::max(max_group, glast[g] - gfirst[g] + 1);
      for (int a = gfirst[g]; a <= glast[g]; ++a) agroup[a] = g;
    }
    // half list (serial nnlpp layout) -> CSR
    std::vector<int> h_off(ngrp + 1, 0), h_idx;
    if (iz == 1) {
      size_t n = 0;
      for (int g = 0; g < ngrp; ++g) {
        int m = nnlpp[n];
        ORAC_REQUIRE(m >= 0, E_ARG, "negative count in nnlpp");
        for (int t = 1; t <= m; ++t) {
          int j = nnlpp[n + t] - 1;
          ORAC_REQUIRE(j >= 0 && j < ngrp, E_ARG, "nnlpp holds a group index out of range");
          h_idx.push_back(j);
        }
        h_off[g + 1] = (int)h_idx.size();
        n += (size_t)m + 1;
      }
    }
    if (h_idx.empty()) h_idx.push_back(0);
    int *d_hoff = B.put_vec("f_hoff", h_off, st, false);
    int *d_hidx = B.put_vec("f_hidx", h_idx, st, false);
    double *d_xpg = B.put("f_xpg", xpg, ngrp, st, false), *d_ypg = B.put("f_ypg", ypg, ngrp, st, false),
           *d_zpg = B.put("f_zpg", zpg, ngrp, st, false);

    // next inner shell list (force.f:278-286, 337-345)
    if (rinn > 1.0e-5) {
      int *cnt = B.get<int>("f_icnt", ngrp + 1), *ioff = B.get<int>("f_ioff", ngrp + 1),
          *itot = B.get<int>("f_itot", 2);
      int tot = 0;
      if (r2inn > 1.0e-5) {
        launch_inner_list(ngrp, iz, d_hoff, d_hidx, d_xpg, d_ypg, d_zpg, co, r2neigh, cnt, nullptr, nullptr, 0, st);
        launch_exclusive_scan(cnt, ioff, ngrp, itot, st);
        ORAC_HIP(hipMemcpyAsync(&tot, itot, sizeof(int), hipMemcpyDeviceToHost, st));
        ORAC_HIP(hipStreamSynchronize(st));
      }
      int64_t need = (int64_t)tot + ngrp;
      ORAC_REQUIRE(nnlpp1 && need <= nnlpp1_capacity, E_CAPACITY,
                   "force: dimensions of NNLPP1 are insufficient (" + std::to_string(need) + " needed)");
      int *out = B.get<int>("f_iout", (size_t)need + 2);
      if (r2inn > 1.0e-5) {
        launch_inner_list(ngrp, iz, d_hoff, d_hidx, d_xpg, d_ypg, d_zpg, co, r2neigh, nullptr, ioff, out, 1, st);
      } else {
        // every group gets an empty record [0]
        launch_zero_u64(reinterpret_cast<unsigned long long *>(out), (size_t)(need + 1) / 2, st);
      }
      ORAC_HIP(hipMemcpyAsync(nnlpp1, out, (size_t)need * sizeof(int), hipMemcpyDeviceToHost, st));
      ORAC_HIP(hipStreamSynchronize(st));
      *ncount1 = (int32_t)need;
    }
    if (!compute) return;

    // atoms and topology
    std::vector<int> tp(nato), ap(nato);
    int nprot = 1;
    for (int a = 0; a < nato; ++a) {
      tp[a] = nbtype[a] - 1;
      ap[a] = atomp[a] - 1;
      ORAC_REQUIRE(tp[a] >= 0 && tp[a] < ma && ap[a] >= 0, E_ARG, "nbtype or atomp out of range");
      nprot = std::max(nprot, atomp[a]);
    }
    PairDesc p;
    std::memset(&p, 0, sizeof p);
    p.natoms = nato;
    p.ngrp = ngrp;
    p.xp = B.put("f_xp", xp0, nato, st, false);
    p.yp = B.put("f_yp", yp0, nato, st, false);
    p.zp = B.put("f_zp", zp0, nato, st, false);
    p.q = B.put("f_q", charge, nato, st, true);
    p.type = B.put_vec("f_type", tp, st, true);
    p.mass = B.put("f_mass", mass, nato, st, true);
    p.gmass = B.put("f_gmass", gmass, nato, st, true);
    p.atomp = B.put_vec("f_atomp", ap, st, true);
    p.xpcm = B.put("f_xpcm", xpcm, nprot, st, false);
    p.ypcm = B.put("f_ypcm", ypcm, nprot, st, false);
    p.zpcm = B.put("f_zpcm", zpcm, nprot, st, false);
    p.gfirst = B.put_vec("f_gfirst", gfirst, st, true);
    p.glast = B.put_vec("f_glast", glast, st, true);
    p.gss = B.put_vec("f_gss", gss, st, true);
    p.xpg = d_xpg;
    p.ypg = d_ypg;
    p.zpg = d_zpg;
    p.ntypes = ma;
    const size_t npk = (size_t)ma * (ma + 1) / 2;
    if (ma <= 256) {
      std::vector<double2> dense((size_t)ma * ma);
      for (int i = 0; i < ma; ++i)
        for (int j = 0; j < ma; ++j) {
          int lij = type[i + (size_t)ma * j] - 1;
          ORAC_REQUIRE(lij >= 0 && (size_t)lij < npk, E_ARG, "type table index out of range");
          dense[(size_t)i * ma + j] = make_double2(ecc12[lij], ecc6[lij]);
        }
      p.lj_dense = B.put_vec("f_ljd", dense, st, true);
    } else {
      std::vector<int> tt((size_t)ma * ma);
      for (size_t k = 0; k < tt.size(); ++k) {
        tt[k] = type[k] - 1;
        ORAC_REQUIRE(tt[k] >= 0 && (size_t)tt[k] < npk, E_ARG, "type table index out of range");
      }
      p.type_table = B.put_vec("f_tt", tt, st, true);
      p.ecc12 = B.put("f_e12", ecc12, npk, st, true);
      p.ecc6 = B.put("f_e6", ecc6, npk, st, true);
    }
    // exclusions (symmetric), rebuilt only when mapnl changes
    {
      size_t mlen = mapnl_length(mapnl, nato);
      auto &c = B.cache["f_mapnl_src"];
      size_t nb = mlen * sizeof(int32_t);
      if (!(c.size() == nb && std::memcmp(c.data(), mapnl, nb) == 0) || !B.bufs.count("f_exoff")) {
        c.assign(reinterpret_cast<const char *>(mapnl), reinterpret_cast<const char *>(mapnl) + nb);
        std::vector<int> eoff, eidx;
        std::vector<unsigned char> inter;
        exclusions_from_mapnl(mapnl, nato, agroup, eoff, eidx, inter);
        B.put_vec("f_exoff", eoff, st, false);
        B.put_vec("f_exidx", eidx, st, false);
        B.put_vec("f_exinter", inter, st, false);
      }
      p.ex_off = B.get<int>("f_exoff", nato + 1);
      p.ex_idx = static_cast<int *>(B.bufs["f_exidx"]->p);
      p.ex_inter = B.get<unsigned char>("f_exinter", nato);
    }
    int heavy = 0;
    for (int a = 0; a < nato && !heavy; ++a) heavy = !(mass[a] < massinfty);
    p.any_heavy = heavy;
    p.massinfty = massinfty;
    // full list from the half list
    int *incount = B.get<int>("f_incount", ngrp + 1), *cursor = B.get<int>("f_cursor", ngrp + 1),
        *f_off = B.get<int>("f_foff", ngrp + 1), *ftot = B.get<int>("f_ftot", 2);
    int *f_idx = B.get<int>("f_fidx", 2 * (size_t)h_off[ngrp] + 1);
    launch_half_to_full(ngrp, d_hoff, d_hidx, incount, f_off, f_idx, cursor, ftot, st);
    p.nb_off = f_off;
    p.nb_idx = f_idx;
    for (int k = 0; k < 9; ++k) p.co[k] = co[k];
    p.alpha = alphal;
    p.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
    p.r2inn = r2inn;
    p.r2out = r2out;
    p.rinn0 = rinn0;
    p.r2inn0 = r2inn0;
    p.rout0 = rout0;
    p.r2out0 = r2out0;
    p.arsout1 = rout0 - 3.0 * rout;
    p.arsout2 = rtolout * rtolout * rtolout;
    p.arsinn1 = rinn0 - 3.0 * rinn;
    p.arsinn2 = *rtolinn * *rtolinn * *rtolinn;
    p.lskip_ewald = lskip_ewald ? 1 : 0;
    p.ewald = ewald ? 1 : 0;
    p.fx = B.get<double>("f_fx", nato);
    p.fy = B.get<double>("f_fy", nato);
    p.fz = B.get<double>("f_fz", nato);
    p.nblk = std::max(1, (ngrp + PAIR_WAVES - 1) / PAIR_WAVES);
    p.part = B.get<double>("f_part", (size_t)p.nblk * 16);
    p.apos = B.get<double4>("f_apos", nato);
    p.ainfo = B.get<int2>("f_ainfo", nato);
    PairDesc *dd = B.put("f_desc", &p, 1, st, false);
    double *dout = B.get<double>("f_out", 16);
    if (ngrp > 0) {
      launch_pairs<double>(dd, p, 1, max_group, st);
      launch_pairs_finalize(dd, 1, dout, st);
    } else {
      launch_zero_f64(dout, 16, st);
    }
    ORAC_HIP(hipGetLastError());
    std::vector<double> fx(nato), fy(nato), fz(nato);
    double out[16];
    if (nato) {
      ORAC_HIP(hipMemcpyAsync(fx.data(), p.fx, nato * sizeof(double), hipMemcpyDeviceToHost, st));
      ORAC_HIP(hipMemcpyAsync(fy.data(), p.fy, nato * sizeof(double), hipMemcpyDeviceToHost, st));
      ORAC_HIP(hipMemcpyAsync(fz.data(), p.fz, nato * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    ORAC_HIP(hipMemcpyAsync(out, dout, sizeof out, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    for (int a = 0; a < nato; ++a) {
      fpx[a] = fpx[a] + fx[a];
      fpy[a] = fpy[a] + fy[a];
      fpz[a] = fpz[a] + fz[a];
    }
    // typeij 1 = slt-slt, 2 = slt-slv, 3 = slv-slv (force.f:849-854)
    *uconf_slt = out[0];
    *uconf_ss = out[1];
    *uconf_slv = out[2];
    *ucoul_slt = out[3];
    *ucoul_ss = out[4];
    *ucoul_slv = out[5];
    const double *s = out + 6;  // st row-major
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) stress[r + 3 * c] = stress[r + 3 * c] + s[3 * r + k] * co[c + 3 * k];
  });
}

// ---------------------------------------------------------------- corrections

static void run_corr_pairs(orac_hip_ctx *ctx, CorrDesc &h, const int32_t *list, int nlist, int natoms,
                           const std::string &pfx, double e_out[4], double *fpx, double *fpy, double *fpz) {
  hipStream_t st = ctx->stream;
  BufSet &B = ctx->B;
  std::vector<int> ia(std::max(1, nlist)), ib(std::max(1, nlist));
  std::vector<std::vector<int>> per(natoms);
  for (int p = 0; p < nlist; ++p) {
    ia[p] = list[2 * p] - 1;
    ib[p] = list[2 * p + 1] - 1;
    ORAC_REQUIRE(ia[p] >= 0 && ia[p] < natoms && ib[p] >= 0 && ib[p] < natoms, E_ARG, "pair list out of range");
    per[ia[p]].push_back(p + 1);
    per[ib[p]].push_back(-(p + 1));
  }
  std::vector<int> goff(natoms + 1, 0), gp;
  for (int a = 0; a < natoms; ++a) {
    for (int v : per[a]) gp.push_back(v);
    goff[a + 1] = (int)gp.size();
  }
  if (gp.empty()) gp.push_back(0);
  h.npairs = nlist;
  h.natoms = natoms;
  h.ia = B.put_vec(pfx + "ia", ia, st, false);
  h.ib = B.put_vec(pfx + "ib", ib, st, false);
  h.g_off = B.put_vec(pfx + "goff", goff, st, false);
  h.g_pair = B.put_vec(pfx + "gp", gp, st, false);
  h.pf = B.get<double>(pfx + "pf", 3 * (size_t)std::max(1, nlist));
  h.pe = B.get<double>(pfx + "pe", 4 * (size_t)std::max(1, nlist));
  h.fx = B.get<double>(pfx + "fx", natoms);
  h.fy = B.get<double>(pfx + "fy", natoms);
  h.fz = B.get<double>(pfx + "fz", natoms);
  launch_zero_f64(h.fx, natoms, st);
  launch_zero_f64(h.fy, natoms, st);
  launch_zero_f64(h.fz, natoms, st);
  CorrDesc *dd = B.put(pfx + "desc", &h, 1, st, false);
  double *dout = B.get<double>(pfx + "out", 4);
  launch_corr(h, dd, st);
  launch_corr_energy(dd, dout, st);
  ORAC_HIP(hipGetLastError());
  std::vector<double> fx(natoms), fy(natoms), fz(natoms);
  if (h.iz == 1 && natoms) {
    ORAC_HIP(hipMemcpyAsync(fx.data(), h.fx, natoms * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fy.data(), h.fy, natoms * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fz.data(), h.fz, natoms * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  ORAC_HIP(hipMemcpyAsync(e_out, dout, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
  ORAC_HIP(hipStreamSynchronize(st));
  if (h.iz == 1)
    for (int a = 0; a < natoms; ++a) {
      fpx[a] = fpx[a] + fx[a];
      fpy[a] = fpy[a] + fy[a];
      fpz[a] = fpz[a] + fz[a];
    }
}

extern "C" int orac_hip_ferrf(orac_hip_ctx *ctx, const int32_t *ss_index, double alphal, const double *charge,
                              double fudge, const double *x0, const double *y0, const double *z0, int32_t iz,
                              const int32_t *list, int32_t nlist, double *fscnstr_slt, double *fscnstr_slv,
                              double *fpx, double *fpy, double *fpz, int32_t erf_corr, int32_t nbonds_added,
                              int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(!erf_corr, E_UNSUPPORTED, "ferrf: the ERF_CORR spline correction is not supported");
    ORAC_REQUIRE(ss_index && charge && x0 && y0 && z0 && fscnstr_slt && fscnstr_slv, E_ARG, "null argument");
    ORAC_REQUIRE(mb >= 0 && nlist >= 0 && nbonds_added >= 0 && nbonds_added <= nlist, E_ARG, "bad sizes");
    int n = nlist - nbonds_added;
    ORAC_REQUIRE(n == 0 || list, E_ARG, "null pair list");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    CorrDesc h;
    std::memset(&h, 0, sizeof h);
    h.x = B.put("c_x", x0, mb, st, false);
    h.y = B.put("c_y", y0, mb, st, false);
    h.z = B.put("c_z", z0, mb, st, false);
    h.q = B.put("c_q", charge, mb, st, true);
    h.ss = B.put("c_ss", ss_index, mb, st, true);
    h.alpha = alphal;
    h.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
    h.fudge = fudge;
    h.iz = iz == 1 ? 1 : 0;
    h.kind = 0;
    double e[4];
    run_corr_pairs(ctx, h, list, n, mb, "c_", e, fpx, fpy, fpz);
    *fscnstr_slt = e[0];
    *fscnstr_slv = e[1];
  });
}

extern "C" int orac_hip_cself(orac_hip_ctx *ctx, const int32_t *ss_index, int32_t nato, double alphal, double kcut,
                              const double *charge, double *self_slt, double *self_slv) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(ss_index && charge && self_slt && self_slv && nato >= 0, E_ARG, "bad argument");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    double *q = B.put("cs_q", charge, nato, st, false);
    int *ss = B.put("cs_ss", ss_index, nato, st, false);
    double *dout = B.get<double>("cs_out", 2);
    launch_cself(nato, q, ss, dout, st);
    double s[2];
    ORAC_HIP(hipMemcpyAsync(s, dout, sizeof s, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    // cself.f:51-56
    const double a1 = 0.2548296, a2 = -0.28449674, a3 = 1.4214137, a4 = -1.453152, a5 = 1.0614054, qp = 0.3275911;
    const double pi = 4.0 * std::atan(1.0);
    for (int n = 0; n < 2; ++n) {
      double alphar = kcut / (2.0 * alphal);
      double qt = 1.0 / (1.0 + qp * alphar);
      double expcst = std::exp(-(alphar * alphar));
      double erfcst = ((((a5 * qt + a4) * qt + a3) * qt + a2) * qt + a1) * qt * expcst;
      s[n] = -s[n] * alphal / std::sqrt(pi) * (1.0 - erfcst);
    }
    *self_slt = s[0];
    *self_slv = s[1];
  });
}

extern "C" int orac_hip_fnbgrp(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                               const double *zp0, const double *charge, const int32_t *type, const double *ecc12,
                               const double *ecc6, int32_t ewald, double alphal, const int32_t *ingrp,
                               int32_t ingrpp, double *fpx, double *fpy, double *fpz, double *uconf_slt,
                               double *uconf_slv, double *ucoul_slt, double *ucoul_slv, int32_t nato) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(uconf_slt && uconf_slv && ucoul_slt && ucoul_slv, E_ARG, "null output");
    *uconf_slt = *uconf_slv = *ucoul_slt = *ucoul_slv = 0.0;
    if (ingrpp == 0) return;
    ORAC_REQUIRE(ss_index && xp0 && yp0 && zp0 && charge && type && ecc12 && ecc6 && ingrp && nato > 0, E_ARG,
                 "null argument");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    int tmax = 1;
    for (int a = 0; a < nato; ++a) {
      ORAC_REQUIRE(type[a] >= 1, E_ARG, "fnbgrp: atom types are 1-based");
      ORAC_REQUIRE(ss_index[a] == 1 || ss_index[a] == 2, E_ARG, "ss_index must be 1 or 2");
      tmax = std::max(tmax, type[a]);
    }
    size_t npk = (size_t)tmax * (tmax + 1) / 2;
    CorrDesc h;
    std::memset(&h, 0, sizeof h);
    h.xp = B.put("g_x", xp0, nato, st, false);
    h.yp = B.put("g_y", yp0, nato, st, false);
    h.zp = B.put("g_z", zp0, nato, st, false);
    h.q = B.put("g_q", charge, nato, st, true);
    h.ss = B.put("g_ss", ss_index, nato, st, true);
    h.type = B.put("g_type", type, nato, st, true);
    h.ecc12 = B.put("g_e12", ecc12, npk, st, true);
    h.ecc6 = B.put("g_e6", ecc6, npk, st, true);
    h.alpha = alphal;
    h.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
    h.iz = 1;
    h.kind = 1;
    h.ewald = ewald ? 1 : 0;
    double e[4];
    run_corr_pairs(ctx, h, ingrp, ingrpp, nato, "g_", e, fpx, fpy, fpz);
    *ucoul_slt = e[0];
    *ucoul_slv = e[1];
    *uconf_slt = e[2];
    *uconf_slv = e[3];
  });
}

// ================================================================ resident systems

namespace orac {

static System &get_sys(orac_hip_ctx *ctx, int32_t sys) {
  ORAC_REQUIRE(ctx, E_ARG, "null context");
  auto it = ctx->systems.find(sys);
  ORAC_REQUIRE(it != ctx->systems.end(), E_ARG, "no such system");
  return *it->second;
}

static void system_build(orac_hip_ctx *ctx, System &S, const orac_hip_system_desc &d) {
  hipStream_t st = ctx->stream;
  BufSet &B = S.B;
  const int n = d.natoms, ng = d.ngrp, ma = d.ntypes;
  ORAC_REQUIRE(n > 0 && ng > 0 && ma > 0 && d.nmol > 0, E_ARG, "system sizes must be positive");
  ORAC_REQUIRE(d.charge && d.mass && d.nbtype && d.type_table && d.ecc12 && d.ecc6 && d.grppt && d.atomp &&
                   d.ss_index && d.mapnl && d.co,
               E_ARG, "null array in system description");
  ORAC_REQUIRE(d.precision == 0 || d.precision == 1, E_ARG, "precision must be 0 or 1");
  S.natoms = n;
  S.ngrp = ng;
  S.nmol = d.nmol;
  S.ntypes = ma;
  S.precision = d.precision;
  S.alpha = d.alpha;
  S.rcut = d.rcut;
  S.rtol = d.rtol;
  S.rnei = d.rnei;
  for (int k = 0; k < 9; ++k) S.co[k] = d.co[k];
  // oc = co^-1, volume = det(co) * 8 (mtsmd.f:1569-1570)
  const double *c = d.co;
  double det = c[0] * (c[4] * c[8] - c[7] * c[5]) - c[3] * (c[1] * c[8] - c[7] * c[2]) + c[6] * (c[1] * c[5] - c[4] * c[2]);
  ORAC_REQUIRE(det > 0.0, E_ARG, "cell matrix must have a positive determinant");
  S.oc[0] = (c[4] * c[8] - c[7] * c[5]) / det;
  S.oc[3] = -(c[3] * c[8] - c[6] * c[5]) / det;
  S.oc[6] = (c[3] * c[7] - c[6] * c[4]) / det;
  S.oc[1] = -(c[1] * c[8] - c[7] * c[2]) / det;
  S.oc[4] = (c[0] * c[8] - c[6] * c[2]) / det;
  S.oc[7] = -(c[0] * c[7] - c[6] * c[1]) / det;
  S.oc[2] = (c[1] * c[5] - c[4] * c[2]) / det;
  S.oc[5] = -(c[0] * c[5] - c[3] * c[2]) / det;
  S.oc[8] = (c[0] * c[4] - c[3] * c[1]) / det;
  S.volume = det * 8.0;

  // groups, molecules, per-atom topology
  std::vector<int> gfirst(ng), glast(ng), gss(ng), agroup(n, -1), tp(n), ap(n);
  std::vector<double> gmass(n);
  for (int g = 0; g < ng; ++g) {
    gfirst[g] = d.grppt[2 * g] - 1;
    glast[g] = d.grppt[2 * g + 1] - 1;
    ORAC_REQUIRE(gfirst[g] >= 0 && glast[g] >= gfirst[g] && glast[g] < n, E_ARG, "bad grppt");
    gss[g] = d.ss_index[gfirst[g]];
    S.max_group = std::max(S.max_group, glast[g] - gfirst[g] + 1);
    double m = 0.0;
    for (int a = gfirst[g]; a <= glast[g]; ++a) {
      agroup[a] = g;
      m += d.mass[a];
    }
    for (int a = gfirst[g]; a <= glast[g]; ++a) gmass[a] = d.mass[a] / m;  // redmss.f
  }
  S.mfirst.assign(d.nmol, -1);
  S.mlast.assign(d.nmol, -1);
  for (int a = 0; a < n; ++a) {
    ORAC_REQUIRE(agroup[a] >= 0, E_ARG, "an atom belongs to no group");
    ORAC_REQUIRE(d.ss_index[a] == 1 || d.ss_index[a] == 2, E_ARG, "ss_index must be 1 (solute) or 2 (solvent)");
    tp[a] = d.nbtype[a] - 1;
    ap[a] = d.atomp[a] - 1;
    ORAC_REQUIRE(tp[a] >= 0 && tp[a] < ma && ap[a] >= 0 && ap[a] < d.nmol, E_ARG, "nbtype or atomp out of range");
    if (S.mfirst[ap[a]] < 0) S.mfirst[ap[a]] = a;
    ORAC_REQUIRE(S.mlast[ap[a]] < 0 || S.mlast[ap[a]] == a - 1, E_ARG, "molecules must be contiguous atom ranges");
    S.mlast[ap[a]] = a;
  }
  S.xp = B.get<double>("xp", n);
  S.yp = B.get<double>("yp", n);
  S.zp = B.get<double>("zp", n);
  S.fx = B.get<double>("fx", n);
  S.fy = B.get<double>("fy", n);
  S.fz = B.get<double>("fz", n);
  launch_zero_f64(S.xp, n, st);
  launch_zero_f64(S.yp, n, st);
  launch_zero_f64(S.zp, n, st);
  double *q = B.put("q", d.charge, n, st, false);
  double *mass = B.put("mass", d.mass, n, st, false);
  int *d_ap = B.put_vec("atomp", ap, st, false);
  int *d_ss = B.put("ss", d.ss_index, n, st, false);
  double *xpcm = B.get<double>("xpcm", d.nmol), *ypcm = B.get<double>("ypcm", d.nmol),
         *zpcm = B.get<double>("zpcm", d.nmol);
  double *xpg = B.get<double>("xpg", ng), *ypg = B.get<double>("ypg", ng), *zpg = B.get<double>("zpg", ng);
  B.put_vec("mfirst", S.mfirst, st, false);
  B.put_vec("mlast", S.mlast, st, false);

  // ---- direct space
  PairDesc &p = S.pair_h;
  std::memset(&p, 0, sizeof p);
  p.natoms = n;
  p.ngrp = ng;
  p.xp = S.xp;
  p.yp = S.yp;
  p.zp = S.zp;
  p.q = q;
  p.type = B.put_vec("type", tp, st, false);
  p.mass = mass;
  p.gmass = B.put_vec("gmass", gmass, st, false);
  p.atomp = d_ap;
  p.xpcm = xpcm;
  p.ypcm = ypcm;
  p.zpcm = zpcm;
  p.gfirst = B.put_vec("gfirst", gfirst, st, false);
  p.glast = B.put_vec("glast", glast, st, false);
  p.gss = B.put_vec("gss", gss, st, false);
  p.xpg = xpg;
  p.ypg = ypg;
  p.zpg = zpg;
  p.ntypes = ma;
  const size_t npk = (size_t)ma * (ma + 1) / 2;
  if (ma <= 256) {
    std::vector<double2> dense((size_t)ma * ma);
    for (int i = 0; i < ma; ++i)
      for (int j = 0; j < ma; ++j) {
        int lij = d.type_table[i + (size_t)ma * j] - 1;
        ORAC_REQUIRE(lij >= 0 && (size_t)lij < npk, E_ARG, "type table index out of range");
        dense[(size_t)i * ma + j] = make_double2(d.ecc12[lij], d.ecc6[lij]);
      }
    p.lj_dense = B.put_vec("ljd", dense, st, false);
  } else {
    std::vector<int> tt((size_t)ma * ma);
    for (size_t k = 0; k < tt.size(); ++k) {
      tt[k] = d.type_table[k] - 1;
      ORAC_REQUIRE(tt[k] >= 0 && (size_t)tt[k] < npk, E_ARG, "type table index out of range");
    }
    p.type_table = B.put_vec("tt", tt, st, false);
    p.ecc12 = B.put("e12", d.ecc12, npk, st, false);
    p.ecc6 = B.put("e6", d.ecc6, npk, st, false);
  }
  {
    std::vector<int> eoff, eidx;
    std::vector<unsigned char> inter;
    exclusions_from_mapnl(d.mapnl, n, agroup, eoff, eidx, inter);
    p.ex_off = B.put_vec("exoff", eoff, st, false);
    p.ex_idx = B.put_vec("exidx", eidx, st, false);
    p.ex_inter = B.put_vec("exinter", inter, st, false);
  }
  p.any_heavy = 0;
  p.massinfty = 1.0e20;
  for (int k = 0; k < 9; ++k) p.co[k] = d.co[k];
  p.alpha = d.alpha;
  p.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
  // single shell: rinn = 0, rout = rcut switched over rtol (mts_forces.f:88-93)
  const double rinn = 0.0, rout = d.rcut, rtolout = d.rtol;
  p.r2inn = 0.0;
  p.r2out = rout * rout;
  p.rinn0 = rinn;
  p.r2inn0 = 0.0;
  p.rout0 = rout + rtolout;
  p.r2out0 = p.rout0 * p.rout0;
  p.arsout1 = p.rout0 - 3.0 * rout;
  p.arsout2 = rtolout * rtolout * rtolout;
  p.arsinn1 = 0.0;
  p.arsinn2 = 0.0;
  p.lskip_ewald = 0;
  p.ewald = 1;
  p.fx = S.fx;
  p.fy = S.fy;
  p.fz = S.fz;
  p.nblk = (ng + PAIR_WAVES - 1) / PAIR_WAVES;
  p.part = B.get<double>("ppart", (size_t)p.nblk * 16);
  p.apos = B.get<double4>("apos", n);
  p.ainfo = B.get<int2>("ainfo", n);

  // ---- linked cells and the full list
  {
    const double rl = d.rcut + d.rtol + d.rnei;
    std::vector<int32_t> nind(2), ii(2 * 4096), jj(2 * 4096), kk(2 * 4096);
    ORAC_REQUIRE(d.ncx > 0 && d.ncy > 0 && d.ncz > 0, E_ARG, "cell counts must be positive");
    int ncx = d.ncx, ncy = d.ncy, ncz = d.ncz;
    if (!lc_index(4096, ncx, ncy, ncz, nind.data(), ii.data(), jj.data(), kk.data(), rl, d.co)) {
      // box too small for linked cells at this cutoff: one cell, every pair tested
      // under the minimum image (the reference's non-linked-cell update)
      ncx = ncy = ncz = 1;
      ORAC_REQUIRE(lc_index(4096, 1, 1, 1, nind.data(), ii.data(), jj.data(), kk.data(), rl, d.co), E_STATE,
                   "single-cell stencil failed");
    }
    S.cg = make_cellgrid(B, "", ncx, ncy, ncz, nind.data(), ii.data(), jj.data(), kk.data(), rl, d.co, 0, st);
    S.cg.cpos = B.get<double4>("cpos", ng);
    S.ncell = ncx * ncy * ncz;
    double dens = ng / (S.volume);
    double nb = 4.0 / 3.0 * M_PI * rl * rl * rl * dens;
    S.list_cap = (size_t)(ng * (1.6 * nb + 64.0));
    B.get<int>("nbidx", S.list_cap);
    S.list_total = B.get<int>("ltotal", 2);
    launch_zero_u64(reinterpret_cast<unsigned long long *>(S.list_total), 1, st);
  }
  p.nb_off = B.get<int>("nboff", ng + 1);
  p.nb_idx = B.get<int>("nbidx", S.list_cap);

  // ---- reciprocal space
  build_pme_plan(B, "", S.plan, d.nfft1, d.nfft2, d.nfft3, d.order, st);
  PmeDesc &m = S.pme_h;
  m = pme_desc_base(B, "", S.plan, n, S.oc, S.volume, d.alpha, 1.0e9, d.precision == 0, ctx->tiles.lds_elems, st);
  m.xp = S.xp;
  m.yp = S.yp;
  m.zp = S.zp;
  m.co = B.put("co", d.co, 9, st, false);
  m.q = q;
  m.fixscale = fix_scale_for(d.charge, n);
  m.fixinv = 1.0 / m.fixscale;
  m.fx = S.fx;
  m.fy = S.fy;
  m.fz = S.fz;
  m.accumulate = 1;
  m.atomp = d_ap;
  m.xpcm = xpcm;
  m.ypcm = ypcm;
  m.zpcm = zpcm;

  // ---- per-step reciprocal corrections (ferrf, iz = 1)
  CorrDesc &h = S.corr_h;
  std::memset(&h, 0, sizeof h);
  {
    int nc = d.ncorr;
    std::vector<int> ia(std::max(1, nc)), ib(std::max(1, nc));
    std::vector<double> fud(std::max(1, nc), 1.0);
    std::vector<std::vector<int>> per(n);
    for (int k = 0; k < nc; ++k) {
      ia[k] = d.corr_pairs[2 * k] - 1;
      ib[k] = d.corr_pairs[2 * k + 1] - 1;
      ORAC_REQUIRE(ia[k] >= 0 && ia[k] < n && ib[k] >= 0 && ib[k] < n, E_ARG, "correction pair out of range");
      if (d.corr_fudge) fud[k] = d.corr_fudge[k];
      per[ia[k]].push_back(k + 1);
      per[ib[k]].push_back(-(k + 1));
    }
    std::vector<int> goff(n + 1, 0), gp;
    for (int a = 0; a < n; ++a) {
      for (int v : per[a]) gp.push_back(v);
      goff[a + 1] = (int)gp.size();
    }
    if (gp.empty()) gp.push_back(0);
    h.npairs = nc;
    h.natoms = n;
    h.ia = B.put_vec("cia", ia, st, false);
    h.ib = B.put_vec("cib", ib, st, false);
    h.fudge_p = B.put_vec("cfud", fud, st, false);
    h.g_off = B.put_vec("cgoff", goff, st, false);
    h.g_pair = B.put_vec("cgp", gp, st, false);
    h.pf = B.get<double>("cpf", 3 * (size_t)std::max(1, nc));
    h.pe = B.get<double>("cpe", 4 * (size_t)std::max(1, nc));
  }
  h.xp = S.xp;
  h.yp = S.yp;
  h.zp = S.zp;
  for (int k = 0; k < 9; ++k) h.co[k] = d.co[k];
  h.q = q;
  h.ss = d_ss;
  h.alpha = d.alpha;
  h.twrtpi = p.twrtpi;
  h.fudge = 1.0;
  h.iz = 1;
  h.kind = 0;
  h.fx = S.fx;
  h.fy = S.fy;
  h.fz = S.fz;

  // self energy (cself.f, kcut = 1e9): constant for fixed charges
  {
    double s[2] = {0.0, 0.0};
    for (int a = 0; a < n; ++a) s[d.ss_index[a] - 1] += d.charge[a] * d.charge[a];
    const double pi = 4.0 * std::atan(1.0);
    for (int k = 0; k < 2; ++k) S.self_e[k] = -s[k] * d.alpha / std::sqrt(pi);
  }
  S.out = B.get<double>("out", 48);
  launch_zero_f64(S.out, 48, st);
  S.pme_d = B.put("pmedesc", &m, 1, st, false);
  S.pair_d = B.put("pairdesc", &p, 1, st, false);
  S.corr_d = B.put("corrdesc", &h, 1, st, false);
  ORAC_HIP(hipStreamSynchronize(st));
}

static void system_eval(orac_hip_ctx *ctx, System &S, bool rebuild) {
  hipStream_t st = ctx->stream;
  BufSet &B = S.B;
  auto mark = [&](int k) {
    if (S.profile) ORAC_HIP(hipEventRecord(S.ev[k], st));
  };
  const PairDesc &p = S.pair_h;
  mark(ST_LIST);
  launch_mol_com(S.nmol, B.get<int>("mfirst", S.nmol), B.get<int>("mlast", S.nmol), p.mass, S.xp, S.yp, S.zp,
                 const_cast<double *>(p.xpcm), const_cast<double *>(p.ypcm), const_cast<double *>(p.zpcm), st);
  launch_group_centres(S.ngrp, p.gfirst, p.glast, p.gmass, S.xp, S.yp, S.zp, const_cast<double *>(p.xpg),
                       const_cast<double *>(p.ypg), const_cast<double *>(p.zpg), st);
  if (rebuild || !S.have_list) {
    int *cell_of = B.get<int>("cellof", S.ngrp), *cnt = B.get<int>("ccnt", S.ncell),
        *start = B.get<int>("cstart", S.ncell), *cur = B.get<int>("ccur", S.ncell),
        *items = B.get<int>("citems", S.ngrp), *count = B.get<int>("lcount", S.ngrp + 1);
    int *off = const_cast<int *>(p.nb_off), *idx = const_cast<int *>(p.nb_idx);
    launch_cell_build(S.ngrp, p.xpg, p.ypg, p.zpg, S.cg, cell_of, cnt, start, cur, items, st);
    launch_list_count(S.ngrp, p.xpg, p.ypg, p.zpg, S.cg, cell_of, start, cnt, items, count, st);
    launch_exclusive_scan(count, off, S.ngrp, S.list_total, st);
    launch_list_guard(S.ngrp, off, S.list_total, (int)S.list_cap, st);
    launch_list_write(S.ngrp, p.xpg, p.ypg, p.zpg, S.cg, cell_of, start, cnt, items, off, idx, 0, st);
    S.have_list = true;
  }
  mark(ST_CENTRES);
  mark(ST_PAIRS);
  if (S.precision == 0)
    launch_pairs<double>(S.pair_d, p, 1, S.max_group, st);
  else
    launch_pairs<float>(S.pair_d, p, 1, S.max_group, st);
  launch_pairs_finalize(S.pair_d, 1, S.out + 20, st);
  mark(ST_PME);
  if (S.precision == 0)
    launch_pme<double>(S.pme_d, S.pme_h, 1, S.out, st, ctx->tiles);
  else
    launch_pme<float>(S.pme_d, S.pme_h, 1, S.out, st, ctx->tiles);
  mark(ST_CORR);
  if (S.corr_h.npairs > 0) {
    launch_corr(S.corr_h, S.corr_d, st);
    launch_corr_energy(S.corr_d, S.out + 36, st);
  }
  mark(ST_END);
  ORAC_HIP(hipGetLastError());
}

}  // namespace orac

extern "C" int orac_hip_system_create(orac_hip_ctx *ctx, const orac_hip_system_desc *desc, int32_t *sys) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(desc && sys, E_ARG, "null argument");
    auto S = std::make_unique<System>();
    system_build(ctx, *S, *desc);
    for (auto &e : S->ev) ORAC_HIP(hipEventCreate(&e));
    int id = ctx->next_sys++;
    ctx->systems[id] = std::move(S);
    *sys = id;
  });
}

extern "C" int orac_hip_system_destroy(orac_hip_ctx *ctx, int32_t sys) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &e : S.ev)
      if (e) ORAC_HIP(hipEventDestroy(e));
    ctx->systems.erase(sys);
  });
}

extern "C" int orac_hip_system_set_coords(orac_hip_ctx *ctx, int32_t sys, const double *xp, const double *yp,
                                          const double *zp) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_REQUIRE(xp && yp && zp, E_ARG, "null coordinates");
    size_t nb = (size_t)S.natoms * sizeof(double);
    ORAC_HIP(hipMemcpyAsync(S.xp, xp, nb, hipMemcpyHostToDevice, ctx->stream));
    ORAC_HIP(hipMemcpyAsync(S.yp, yp, nb, hipMemcpyHostToDevice, ctx->stream));
    ORAC_HIP(hipMemcpyAsync(S.zp, zp, nb, hipMemcpyHostToDevice, ctx->stream));
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int orac_hip_system_device_ptrs(orac_hip_ctx *ctx, int32_t sys, double **xp, double **yp, double **zp,
                                           double **fx, double **fy, double **fz) {
  return guarded([&] {
    System &S = get_sys(ctx, sys);
    if (xp) *xp = S.xp;
    if (yp) *yp = S.yp;
    if (zp) *zp = S.zp;
    if (fx) *fx = S.fx;
    if (fy) *fy = S.fy;
    if (fz) *fz = S.fz;
  });
}

extern "C" int orac_hip_system_eval(orac_hip_ctx *ctx, int32_t sys, int32_t rebuild_list) {
  return guarded([&] {
    set_device(ctx);
    system_eval(ctx, get_sys(ctx, sys), rebuild_list != 0);
  });
}

extern "C" int orac_hip_system_profile(orac_hip_ctx *ctx, int32_t sys, int32_t enable) {
  return guarded([&] { get_sys(ctx, sys).profile = enable != 0; });
}

extern "C" int orac_hip_system_stage_ms(orac_hip_ctx *ctx, int32_t sys, float *ms, int32_t n) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_REQUIRE(ms && n >= 0, E_ARG, "bad buffer");
    ORAC_REQUIRE(S.profile, E_STATE, "profiling is not enabled for this system");
    ORAC_HIP(hipEventSynchronize(S.ev[ST_END]));
    const int pairs[5][2] = {{ST_LIST, ST_CENTRES}, {ST_CENTRES, ST_PAIRS}, {ST_PAIRS, ST_PME}, {ST_PME, ST_CORR},
                             {ST_CORR, ST_END}};
    for (int k = 0; k < n && k < 5; ++k) ORAC_HIP(hipEventElapsedTime(&ms[k], S.ev[pairs[k][0]], S.ev[pairs[k][1]]));
  });
}

extern "C" int orac_hip_system_results(orac_hip_ctx *ctx, int32_t sys, double *fx, double *fy, double *fz,
                                       double *energies) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    hipStream_t st = ctx->stream;
    double out[48];
    int tot[2];
    ORAC_HIP(hipMemcpyAsync(out, S.out, sizeof out, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(tot, S.list_total, sizeof tot, hipMemcpyDeviceToHost, st));
    size_t nb = (size_t)S.natoms * sizeof(double);
    if (fx) ORAC_HIP(hipMemcpyAsync(fx, S.fx, nb, hipMemcpyDeviceToHost, st));
    if (fy) ORAC_HIP(hipMemcpyAsync(fy, S.fy, nb, hipMemcpyDeviceToHost, st));
    if (fz) ORAC_HIP(hipMemcpyAsync(fz, S.fz, nb, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    ORAC_REQUIRE(tot[1] == 0, E_CAPACITY,
                 "neighbour list capacity exceeded (" + std::to_string(tot[0]) + " entries); results are invalid");
    if (energies) {
      energies[0] = out[0];
      // pairs: uconf[typeij], ucoul[typeij] with typeij 1 slt, 2 slt-slv, 3 slv
      energies[1] = out[20 + 0];
      energies[2] = out[20 + 2];
      energies[3] = out[20 + 1];
      energies[4] = out[20 + 3];
      energies[5] = out[20 + 5];
      energies[6] = out[20 + 4];
      energies[7] = out[36 + 0];
      energies[8] = out[36 + 1];
      // stress: direct st.co^T + (-vir + molecular correction)
      const double *s = out + 26;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double v = 0.0;
          for (int k = 0; k < 3; ++k) v += s[3 * r + k] * S.co[c + 3 * k];
          v = v - out[1 + r + 3 * c] + out[10 + 3 * r + c];
          energies[9 + r + 3 * c] = v;
        }
      energies[18] = (double)tot[0];
      energies[19] = S.self_e[0];
      energies[20] = S.self_e[1];
    }
  });
}
