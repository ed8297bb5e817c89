// api.cpp -- C ABI (include/orac_hip.h) of the gfx950 nonbonded path: the
// context and the plain entry points with the reference's call semantics.
//
// Plain-mode entry points convert the reference's arrays to the device
// layouts, run the kernels and copy results back with the reference's
// accumulate/overwrite semantics.  Static topology arrays are compared with
// the copy last uploaded and sent again only when their bytes change.
// Resident systems and batches are in engine.cpp.
#include <cmath>
#include <mutex>

#include "host.hpp"

using namespace orac;

// ---------------------------------------------------------------- static scratch
//
// The device region behind orac_hip_fft_pme_noalloc.  Each GPU the library
// runs on holds its own copy for the life of the process.  Host code never
// uses the array's name as an address: the device address of each GPU's copy
// is looked up once per device, at context creation with that device
// current, and kept in g_scratch_ptr.
namespace {
constexpr size_t SCRATCH_BYTES = (size_t)32 << 20;  // 32 MiB
constexpr int SCRATCH_MAX_ATOMS = 65536;
constexpr int SCRATCH_MAX_POINTS = 64 * 64 * 64;
constexpr int SCRATCH_MAX_DEVICES = 64;
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// device address of this device's scratch copy (the device must be current)
void *scratch_for_device(int device) {
  ORAC_REQUIRE(device >= 0 && device < SCRATCH_MAX_DEVICES, E_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    ORAC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ORAC_REQUIRE(p != nullptr, E_HIP, "hipGetSymbolAddress returned no address for the scratch region");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

// atomp (1-based, fft_pme.f) to the 0-based molecule index the gather reads
__global__ void __launch_bounds__(256) k_atomp_to_zero_based(int *__restrict__ ap, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ap[i] -= 1;
}
}  // namespace

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
    ORAC_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    c->S.arena = static_cast<char *>(scratch_for_device(device));
    c->S.arena_cap = SCRATCH_BYTES;
    *ctx = c.release();
  });
}

extern "C" int orac_hip_destroy(orac_hip_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    engine_destroy(ctx->engine);
    ctx->engine = nullptr;
    ctx->B.bufs.clear();
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
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

// plain copy of n doubles, one element per thread, for byte-counter calibration
__global__ void __launch_bounds__(256) k_probe_copy(const double *__restrict__ a, double *__restrict__ b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

extern "C" int orac_hip_probe_copy(orac_hip_ctx *ctx, int64_t nbytes) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(nbytes >= 8 && nbytes <= (int64_t)1 << 34, E_ARG, "probe size must be 8 B .. 16 GiB");
    const size_t n = (size_t)nbytes / 8;
    DevBuf a, b;
    a.ensure(n * 8);
    b.ensure(n * 8);
    ORAC_HIP(hipMemsetAsync(a.p, 0, n * 8, ctx->stream));
    hipLaunchKernelGGL(k_probe_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const double *)a.p, (double *)b.p, n);
    ORAC_HIP(hipGetLastError());
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
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
    // mts_forces.f:232-239, in the reference's association: (qp alphal) rinn in
    // the rational term and ((alphal rinn) alphal) rinn in the exponent, so the
    // decision agrees with the reference to the last bit at the 1e-3 threshold
    const double a1 = 0.2548296, a2 = -0.28449674, a3 = 1.4214137, a4 = -1.453152, a5 = 1.0614054;
    const double qp = 0.3275911;
    const double qt = 1.0 / (1.0 + qp * alphal * rinn);
    const double expcst = std::exp(-alphal * rinn * alphal * rinn);
    const double erfcst = ((((a5 * qt + a4) * qt + a3) * qt + a2) * qt + a1) * qt * expcst;
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

static int fft_pme_impl(orac_hip_ctx *ctx, int32_t ntap, const double *xp, const double *yp,
                                const double *zp, const double *xpcm, const double *ypcm, const double *zpcm,
                                const double *pmechg, const double *co, const double *oc, double volume,
                                double alphal, int32_t order, int32_t nfft1, int32_t nfft2, int32_t nfft3,
                                double *eer, double *fpx, double *fpy, double *fpz, double *stressc,
                                const int32_t *atomp, const int32_t *grppt, int32_t pressure, double rkcut,
                        const double *charg0, double *dwrk) {
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
    PmeDesc d = pme_desc_base(B, "p_", ctx->plan, ntap, oc, volume, alphal, rkcut, true, ctx->tuning, st);
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
    double *dout = B.get<double>("p_out", 40);
    d.out = dout;
    PmeDesc descs[2] = {d, d};
    int ndesc = 1;
    if (charg0) {
      // do_pmesh_alchemy.f90: the lambda0 charges on a second grid, energy only
      PmeDesc &e = descs[1];
      e = pme_desc_base(B, "p0_", ctx->plan, ntap, oc, volume, alphal, rkcut, true, ctx->tuning, st);
      e.xp = d.xp;
      e.yp = d.yp;
      e.zp = d.zp;
      e.co = d.co;
      e.q = B.put("p0_q", charg0, ntap, st, true);
      e.fixscale = fix_scale_for(charg0, ntap);
      e.fixinv = 1.0 / e.fixscale;
      e.fx = d.fx;
      e.fy = d.fy;
      e.fz = d.fz;
      e.out = dout + 20;
      e.energy_only = 1;
      ndesc = 2;
    }
    PmeDesc *dd = B.put("p_desc", descs, ndesc, st, false);
    launch_pme<double>(dd, descs, ndesc, st, ctx->tuning);
    ORAC_HIP(hipGetLastError());
    std::vector<double> fx(ntap), fy(ntap), fz(ntap);
    double out[40];
    ORAC_HIP(hipMemcpyAsync(fx.data(), d.fx, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fy.data(), d.fy, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fz.data(), d.fz, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(out, dout, sizeof out, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    *eer = out[0];
    if (charg0) *dwrk = out[0] - out[20];  // do_pmesh_alchemy.f90: dwrk = eer - eer0
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

extern "C" int orac_hip_fft_pme(orac_hip_ctx *ctx, int32_t ntap, const double *xp, const double *yp,
                                const double *zp, const double *xpcm, const double *ypcm, const double *zpcm,
                                const double *pmechg, const double *co, const double *oc, double volume,
                                double alphal, int32_t order, int32_t nfft1, int32_t nfft2, int32_t nfft3,
                                double *eer, double *fpx, double *fpy, double *fpz, double *stressc,
                                const int32_t *atomp, const int32_t *grppt, int32_t pressure, double rkcut) {
  return fft_pme_impl(ctx, ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, co, oc, volume, alphal, order, nfft1, nfft2,
                      nfft3, eer, fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut, nullptr, nullptr);
}

extern "C" int orac_hip_fft_pme_alchemy(orac_hip_ctx *ctx, int32_t ntap, const double *xp, const double *yp,
                                        const double *zp, const double *xpcm, const double *ypcm, const double *zpcm,
                                        const double *pmechg, const double *charg0, const double *co,
                                        const double *oc, double volume, double alphal, int32_t order, int32_t nfft1,
                                        int32_t nfft2, int32_t nfft3, double *eer, double *fpx, double *fpy,
                                        double *fpz, double *stressc, const int32_t *atomp, const int32_t *grppt,
                                        int32_t pressure, double rkcut, double *dwrk) {
  if (!charg0 || !dwrk) return guarded([] { ORAC_REQUIRE(false, E_ARG, "null charg0 or dwrk"); });
  return fft_pme_impl(ctx, ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, co, oc, volume, alphal, order, nfft1, nfft2,
                      nfft3, eer, fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut, charg0, dwrk);
}

// fft_pme without allocation: every per-call device buffer is carved out of
// this device's static scratch region (ctx->S), the forces are accumulated on
// the device into the caller's values, and nothing is allocated on the host
// or the device.  Calls that overlap on one GPU share the region and race.
extern "C" int orac_hip_fft_pme_noalloc(orac_hip_ctx *ctx, int32_t ntap, const double *xp, const double *yp,
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
    ORAC_REQUIRE(ntap <= SCRATCH_MAX_ATOMS && (int64_t)nfft1 * nfft2 * nfft3 <= SCRATCH_MAX_POINTS, E_CAPACITY,
                 "fft_pme_noalloc: at most " + std::to_string(SCRATCH_MAX_ATOMS) + " atoms and " +
                     std::to_string(SCRATCH_MAX_POINTS) + " grid points fit the static scratch region (got " +
                     std::to_string(ntap) + " atoms, " + std::to_string((int64_t)nfft1 * nfft2 * nfft3) +
                     " points)");
    int nprot = 0;
    if (pressure) {
      ORAC_REQUIRE(atomp && xpcm && ypcm && zpcm, E_ARG, "pressure needs atomp and xpcm");
      for (int i = 0; i < ntap; ++i) {
        ORAC_REQUIRE(atomp[i] >= 1, E_ARG, "atomp holds an index below 1");
        nprot = std::max(nprot, atomp[i]);
      }
      ORAC_REQUIRE(nprot <= ntap, E_ARG, "atomp holds an index above the atom count");
    }
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->S;
    B.reset();
    PmeDesc d = pme_desc_base(B, "p_", ctx->plan, ntap, oc, volume, alphal, rkcut, true, ctx->tuning, st);
    d.xp = B.put("p_xp", xp, ntap, st, false);
    d.yp = B.put("p_yp", yp, ntap, st, false);
    d.zp = B.put("p_zp", zp, ntap, st, false);
    d.co = B.put("p_co", co, 9, st, false);
    d.q = B.put("p_q", pmechg, ntap, st, false);
    d.fixscale = fix_scale_for(pmechg, ntap);
    d.fixinv = 1.0 / d.fixscale;
    // the gather adds the reciprocal forces to the caller's (fp += f, fft_pme.f:108-112)
    d.fx = B.put("p_fx", fpx, ntap, st, false);
    d.fy = B.put("p_fy", fpy, ntap, st, false);
    d.fz = B.put("p_fz", fpz, ntap, st, false);
    d.accumulate = 1;
    if (pressure) {
      int *ap = B.put("p_atomp", atomp, ntap, st, false);
      if (ntap > 0) hipLaunchKernelGGL(k_atomp_to_zero_based, dim3((ntap + 255) / 256), dim3(256), 0, st, ap, ntap);
      d.atomp = ap;
      d.xpcm = B.put("p_xpcm", xpcm, nprot, st, false);
      d.ypcm = B.put("p_ypcm", ypcm, nprot, st, false);
      d.zpcm = B.put("p_zpcm", zpcm, nprot, st, false);
    }
    double *dout = B.get<double>("p_out", 40);
    d.out = dout;
    PmeDesc *dd = B.put("p_desc", &d, 1, st, false);
    launch_pme<double>(dd, &d, 1, st, ctx->tuning);
    ORAC_HIP(hipGetLastError());
    double out[40];
    ORAC_HIP(hipMemcpyAsync(fpx, d.fx, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fpy, d.fy, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fpz, d.fz, ntap * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(out, dout, sizeof out, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    *eer = out[0];
    if (pressure) {
      // stressc(r,c) = stressc(r,c) - vir(r,c) + stc(3r+c)   (fft_pme.f:97-105)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) stressc[r + 3 * c] = stressc[r + 3 * c] - out[1 + r + 3 * c] + out[10 + 3 * r + c];
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
    if (nentries) *nentries = natp;
    if (natp == 0) return;
    ListDesc L;
    std::memset(&L, 0, sizeof L);
    L.ngrp = natp;
    L.cg = make_cellgrid(B, "l_", ncx, ncy, ncz, nind, indxi, indxj, indxk, rcut, co, 1, st);
    L.cg.cpos = B.get<double4>("l_cpos", natp);
    const int ncell = ncx * ncy * ncz;
    L.xpg = B.put("l_x", xp, natp, st, false);
    L.ypg = B.put("l_y", yp, natp, st, false);
    L.zpg = B.put("l_z", zp, natp, st, false);
    L.cell_of = B.get<int>("l_cellof", natp);
    L.cell_cnt = B.get<int>("l_cnt", ncell + 1);
    L.cell_start = B.get<int>("l_start", ncell + 1);
    L.cell_cursor = B.get<int>("l_cur", ncell + 1);
    L.cell_items = B.get<int>("l_items", natp);
    L.cell_fill = B.get<int>("l_fill", natp);
    L.count = B.get<int>("l_count", natp + 1);
    L.off = B.get<int>("l_off", natp + 1);
    L.total = B.get<int>("l_total", 4);  // entries, overflow, largest count, sticky overflow
    L.cap = nnlpp_capacity;
    L.mode = 1;
    // pass 1: counts and the capacity check against the caller's array
    ListDesc *dd = B.put("l_desc", &L, 1, st, false);
    launch_list_build(dd, &L, 1, LIST_COUNT, st);
    int tot[2] = {0, 0};
    ORAC_HIP(hipMemcpyAsync(tot, L.total, sizeof tot, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    const int64_t need = (int64_t)tot[0] + natp;
    if (nentries) *nentries = need;
    ORAC_REQUIRE(!tot[1] && need <= nnlpp_capacity, E_CAPACITY,
                 "lc_list: dimensions of NNLPP are insufficient (" + std::to_string(need) + " entries needed)");
    // pass 2: the entries, into a device array of exactly that size
    L.out = B.get<int>("l_out", (size_t)need);
    L.cap = need;
    dd = B.put("l_desc", &L, 1, st, false);
    launch_list_build(dd, &L, 1, LIST_WRITE, st);
    ORAC_HIP(hipGetLastError());
    ORAC_HIP(hipMemcpyAsync(nnlpp, L.out, (size_t)need * sizeof(int), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
  });
}

// ---------------------------------------------------------------- force

// force_alchemy.f's extra inputs (module cpropar state passed explicitly)
struct AlchIn {
  const int32_t *ss_alchemy;
  const double *lam;  // ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0
  int intramol, ljintra, opls;
  const double *rmin, *epsm;
  double *dwrk;
  double qsoft, ljsoft;  // (qshift, ljshift) when shifted
  int shifted;
};

// force_REM's extra inputs: rematom (per atom, nonzero in the segment),
// rem_segkind, rem_factor(3), lskip_force and ljintra (module state in the
// reference, arguments here), and the segment energy it returns
struct RemIn {
  const int32_t *rematom;
  int kind;
  double f3;
  int lskip_force, ljintra;
  double *seg_ene;
};

static int force_impl(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                      const double *zp0, const double *xpg, const double *ypg, const double *zpg, const double *charge,
                      const int32_t *nbtype, const int32_t *type, int32_t ma, int32_t nato, const int32_t *atomg,
                      const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *groupp,
                      const int32_t *atomp, const double *co, const double *ecc12, const double *ecc6, double alphal,
                      const int32_t *mapnl, int32_t ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv,
                      double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx,
                      double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1,
                      int64_t nnlpp1_capacity, double rneigh, double rinn, double rout, double *rtolinn,
                      double rtolout, const double *gmass, int32_t iz, int32_t *ncount1, const double *mass,
                      double massinfty, int32_t lskip_ewald, int32_t ewald, const AlchIn *al,
                      const RemIn *rem = nullptr) {
  return guarded([&] {
    set_device(ctx);
    (void)groupp;
    (void)atomg;
    std::vector<double> q_off;
    if (rem) {
      ORAC_REQUIRE(rem->rematom && rem->seg_ene, E_ARG, "null rematom or seg_ene");
      ORAC_REQUIRE(rem->kind >= 0 && rem->kind <= 2, E_ARG, "rem_segkind must be 0, 1 or 2");
      // force_REM.f:143-157: lskip_force without ljintra computes nothing
      // (the outputs are left as they are); with it the charges are zero
      if (rem->lskip_force && !rem->ljintra) return;
      // force_REM.f:491 and :667 read the Coulomb energy or the distance a
      // previous pair or group left in the routine's locals when lskip_ewald
      // is set: an order-dependent result with no parallel form
      ORAC_REQUIRE(!lskip_ewald && ewald, E_UNSUPPORTED,
                   "force_rem: only the Ewald branch (lskip_ewald = .false., ewald = .true.) is supported; the "
                   "others carry over locals from earlier pairs (force_REM.f:491, :667)");
      *rem->seg_ene = 0.0;
      if (rem->lskip_force) {
        q_off.assign((size_t)std::max(nato, 0), 0.0);
        charge = q_off.data();
      }
    }
    ORAC_REQUIRE(nato >= 0 && ngrp >= 0 && ma >= 1, E_ARG, "bad sizes");
    ORAC_REQUIRE(nato < (1 << 26), E_UNSUPPORTED, "more than 2^26 atoms");
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
    p.nmol = nprot;
    p.gfirst = B.put_vec("f_gfirst", gfirst, st, true);
    p.glast = B.put_vec("f_glast", glast, st, true);
    p.gss = B.put_vec("f_gss", gss, st, true);
    p.xpg = d_xpg;
    p.ypg = d_ypg;
    p.zpg = d_zpg;
    p.ntypes = ma;
    const size_t npk = (size_t)ma * (ma + 1) / 2;
    if (al) {
      // soft-core pairs: per-atom classes replace the LJ types
      ORAC_REQUIRE(al->ss_alchemy && al->lam && al->rmin && al->epsm && al->dwrk, E_ARG, "null alchemy argument");
      AlchTables A = alch_tables(nato, al->ss_alchemy, al->rmin, al->epsm, al->opls);
      double fac[72];
      alch_factors(al->lam, al->intramol != 0, al->ljintra != 0, al->qsoft, al->ljsoft, al->shifted != 0, fac);
      p.alch_shifted = al->shifted ? 1 : 0;
      p.ntypes = A.ncls;
      p.type = B.put_vec("f_atype", A.cls, st, true);
      p.alch_pair = B.put_vec("f_apair", A.pair, st, true);
      p.alch_fac = B.put("f_afac", fac, 72, st, true);
    } else if (ma <= LJ_DENSE_MAX) {
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
    p.split = pair_split_for(ngrp);
    p.nblk = std::max(1, (ngrp + PAIR_WAVES / p.split - 1) / (PAIR_WAVES / p.split));
    p.part = B.get<double>("f_part", (size_t)p.nblk * 16);
    p.sgp = B.get<double4>("f_sgp", std::max(1, ngrp));
    p.sgi = B.get<int4>("f_sgi", std::max(1, ngrp));
    p.sap = B.get<double4>("f_sap", std::max(1, nato));
    p.sapc = B.get<double4>("f_sapc", std::max(1, nato));
    p.sai = B.get<int2>("f_sai", std::max(1, nato));
    p.max_group = max_group;
    if (rem) {
      std::vector<int> rf(nato);
      for (int a = 0; a < nato; ++a) rf[a] = rem->rematom[a] != 0;
      p.rematom = B.put_vec("f_rematom", rf, st, false);
      p.rem_kind = rem->kind;
      p.rem_f3 = rem->f3;
    }
    double *dout = B.get<double>("f_out", 16);
    p.out = dout;
    PairDesc *dd = B.put("f_desc", &p, 1, st, false);
    if (ngrp > 0) {
      launch_pairs<double>(dd, &p, 1, st, ctx->tuning);
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
    // force_alchemy.f:684-688: dwrk += sum of U(lambda) - U(lambda0)
    if (al) *al->dwrk = *al->dwrk + out[15];
    if (rem) *rem->seg_ene = out[15];
  });
}

extern "C" int orac_hip_force_rem(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                                  const double *zp0, const double *xpg, const double *ypg, const double *zpg,
                                  const double *charge, const int32_t *nbtype, const int32_t *type, int32_t ma,
                                  int32_t nato, const int32_t *atomg, const double *xpcm, const double *ypcm,
                                  const double *zpcm, const int32_t *groupp, const int32_t *atomp, const double *co,
                                  const double *ecc12, const double *ecc6, double alphal, const int32_t *mapnl,
                                  int32_t ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv,
                                  double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss,
                                  double *fpx, double *fpy, double *fpz, double *stress, const int32_t *nnlpp,
                                  int32_t *nnlpp1, int64_t nnlpp1_capacity, double rneigh, double rinn, double rout,
                                  double *rtolinn, double rtolout, const double *gmass, int32_t iz, int32_t *ncount1,
                                  const double *mass, double massinfty, double *seg_ene, const int32_t *rematom,
                                  int32_t lskip_ewald, int32_t ewald, int32_t rem_segkind, double rem_factor3,
                                  int32_t lskip_force, int32_t ljintra) {
  const RemIn r{rematom, rem_segkind, rem_factor3, lskip_force, ljintra, seg_ene};
  return force_impl(ctx, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm,
                    zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss,
                    ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, nnlpp1_capacity, rneigh,
                    rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, lskip_ewald, ewald, nullptr,
                    &r);
}

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
  return force_impl(ctx, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm,
                    zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss,
                    ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, nnlpp1_capacity, rneigh,
                    rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, lskip_ewald, ewald, nullptr);
}

extern "C" int orac_hip_force_alchemy(
    orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
    const double *xpg, const double *ypg, const double *zpg, const double *charge, const int32_t *nbtype,
    const int32_t *type, int32_t ma, int32_t nato, const int32_t *atomg, const double *xpcm, const double *ypcm,
    const double *zpcm, const int32_t *groupp, const int32_t *atomp, const double *co, const double *ecc12,
    const double *ecc6, double alphal, const int32_t *mapnl, int32_t ngrp, const int32_t *grppt, double *uconf_slt,
    double *uconf_slv, double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx,
    double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1, int64_t nnlpp1_capacity,
    double rneigh, double rinn, double rout, double *rtolinn, double rtolout, const double *gmass, int32_t iz,
    int32_t *ncount1, const double *mass, double massinfty, const int32_t *ss_alchemy, const double *lam,
    int32_t intramol, int32_t ljintra, int32_t opls, const double *rmin, const double *epsm, double *dwrk,
    int32_t lskip_ewald, int32_t ewald, double qsoft, double ljsoft) {
  AlchIn al{ss_alchemy, lam, intramol, ljintra, opls, rmin, epsm, dwrk, qsoft, ljsoft, 0};
  return force_impl(ctx, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm,
                    zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss,
                    ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, nnlpp1_capacity, rneigh,
                    rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, lskip_ewald, ewald, &al);
}

extern "C" int orac_hip_force_alchemy_shift(
    orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
    const double *xpg, const double *ypg, const double *zpg, const double *charge, const int32_t *nbtype,
    const int32_t *type, int32_t ma, int32_t nato, const int32_t *atomg, const double *xpcm, const double *ypcm,
    const double *zpcm, const int32_t *groupp, const int32_t *atomp, const double *co, const double *ecc12,
    const double *ecc6, double alphal, const int32_t *mapnl, int32_t ngrp, const int32_t *grppt, double *uconf_slt,
    double *uconf_slv, double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx,
    double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1, int64_t nnlpp1_capacity,
    double rneigh, double rinn, double rout, double *rtolinn, double rtolout, const double *gmass, int32_t iz,
    int32_t *ncount1, const double *mass, double massinfty, const int32_t *ss_alchemy, const double *lam,
    int32_t intramol, int32_t ljintra, int32_t opls, const double *rmin, const double *epsm, double *dwrk,
    int32_t lskip_ewald, int32_t ewald, double qshift, double ljshift) {
  AlchIn al{ss_alchemy, lam, intramol, ljintra, opls, rmin, epsm, dwrk, qshift, ljshift, 1};
  return force_impl(ctx, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm,
                    zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss,
                    ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, nnlpp1_capacity, rneigh,
                    rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, lskip_ewald, ewald, &al);
}

// ---------------------------------------------------------------- corrections

// Per-atom gather index of a pair (or term) list: for atom a, the signed slots
// +-(p + 1) of the terms that touch it, in term order (a counting sort).  It is
// rebuilt and uploaded only when the list bytes or the atom count change, so a
// caller that passes the same topology every step pays one memcmp per call.
struct GatherIndex {
  const int *ia = nullptr, *ib = nullptr, *g_off = nullptr, *g_slot = nullptr;
};
static GatherIndex gather_index(orac_hip_ctx *ctx, const std::string &pfx, const int32_t *list, int stride,
                                int arity, int n, int natoms) {
  BufSet &B = ctx->B;
  hipStream_t st = ctx->stream;
  auto &key = B.cache[pfx + "#list"];
  const size_t nb = (size_t)n * stride * sizeof(int32_t);
  GatherIndex g;
  const bool same = key.size() == nb + sizeof(int) && (nb == 0 || std::memcmp(key.data(), list, nb) == 0) &&
                    std::memcmp(key.data() + nb, &natoms, sizeof(int)) == 0 && B.bufs.count(pfx + "goff");
  if (!same) {
    std::vector<int> first(std::max(1, n)), second(std::max(1, n)), cnt(natoms + 1, 0);
    for (int p = 0; p < n; ++p)
      for (int k = 0; k < arity; ++k) {
        const int a = list[(size_t)stride * p + k] - 1;
        ORAC_REQUIRE(a >= 0 && a < natoms, E_ARG, "pair or term list out of range");
        ++cnt[a + 1];
      }
    for (int a = 0; a < natoms; ++a) cnt[a + 1] += cnt[a];
    std::vector<int> slot(std::max<size_t>(1, (size_t)cnt[natoms])), cur(cnt.begin(), cnt.end() - 1);
    for (int p = 0; p < n; ++p) {
      for (int k = 0; k < arity; ++k) {
        const int a = list[(size_t)stride * p + k] - 1;
        // pairs: +(p+1) for the first atom, -(p+1) for the second; terms: arity*p + position
        slot[cur[a]++] = arity == 2 ? (k == 0 ? p + 1 : -(p + 1)) : arity * p + k;
      }
      first[p] = list[(size_t)stride * p] - 1;
      second[p] = arity >= 2 ? list[(size_t)stride * p + 1] - 1 : 0;
    }
    B.put_vec(pfx + "ia", first, st, false);
    B.put_vec(pfx + "ib", second, st, false);
    B.put_vec(pfx + "goff", cnt, st, false);
    B.put_vec(pfx + "gp", slot, st, false);
    key.assign(reinterpret_cast<const char *>(list), reinterpret_cast<const char *>(list) + nb);
    key.insert(key.end(), reinterpret_cast<const char *>(&natoms), reinterpret_cast<const char *>(&natoms) + sizeof(int));
  }
  g.ia = B.get<int>(pfx + "ia", std::max(1, n));
  g.ib = B.get<int>(pfx + "ib", std::max(1, n));
  g.g_off = B.get<int>(pfx + "goff", natoms + 1);
  g.g_slot = B.get<int>(pfx + "gp", 1);
  return g;
}

// One correction launch on pairs list(stride, nlist) (1-based atoms in the
// first two rows): forces added to fp when h.iz == 1, e_out[h.ncol] energies.
static void run_corr_pairs(orac_hip_ctx *ctx, CorrDesc &h, const int32_t *list, int nlist, int natoms,
                           const std::string &pfx, double *e_out, double *fpx, double *fpy, double *fpz,
                           int stride = 2) {
  hipStream_t st = ctx->stream;
  BufSet &B = ctx->B;
  const int ncol = h.ncol ? h.ncol : 4;
  GatherIndex gi = gather_index(ctx, pfx, list, stride, 2, nlist, natoms);
  h.npairs = nlist;
  h.natoms = natoms;
  h.ia = gi.ia;
  h.ib = gi.ib;
  h.g_off = gi.g_off;
  h.g_pair = gi.g_slot;
  h.pf = B.get<double>(pfx + "pf", 3 * (size_t)std::max(1, nlist));
  h.pe = B.get<double>(pfx + "pe", ncol * (size_t)std::max(1, nlist));
  h.fx = B.get<double>(pfx + "fx", natoms);
  h.fy = B.get<double>(pfx + "fy", natoms);
  h.fz = B.get<double>(pfx + "fz", natoms);
  h.write = 1;  // the gather writes every atom's sum (zero where no pair touches it)
  double *dout = B.get<double>(pfx + "out", ncol);
  h.out = dout;
  CorrDesc *dd = B.put(pfx + "desc", &h, 1, st, false);
  launch_corr(dd, &h, 1, st);
  ORAC_HIP(hipGetLastError());
  std::vector<double> fx(natoms), fy(natoms), fz(natoms);
  if (h.iz == 1 && natoms && nlist > 0) {
    ORAC_HIP(hipMemcpyAsync(fx.data(), h.fx, natoms * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fy.data(), h.fy, natoms * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(fz.data(), h.fz, natoms * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  ORAC_HIP(hipMemcpyAsync(e_out, dout, ncol * sizeof(double), hipMemcpyDeviceToHost, st));
  ORAC_HIP(hipStreamSynchronize(st));
  if (h.iz == 1 && nlist > 0)
    for (int a = 0; a < natoms; ++a) {
      fpx[a] = fpx[a] + fx[a];
      fpy[a] = fpy[a] + fy[a];
      fpz[a] = fpz[a] + fz[a];
    }
}

// ERF_CORR: the table rows the pair list reaches (ic of the largest separation),
// uploaded for the correction kernel; a pair below the table's first node is an error
static void erf_table_for(orac_hip_ctx *ctx, CorrDesc &h, const std::string &pfx, const double *tab, double delew,
                          double rlew, const int32_t *list, int n, int mb, const double *x0, const double *y0,
                          const double *z0, int stride = 2) {
  ORAC_REQUIRE(tab && delew > 0.0, E_ARG, "ERF_CORR needs its table and a positive spacing");
  int icmax = 1;
  for (int p = 0; p < n; ++p) {
    const int a = list[(size_t)stride * p] - 1, b = list[(size_t)stride * p + 1] - 1;
    ORAC_REQUIRE(a >= 0 && a < mb && b >= 0 && b < mb, E_ARG, "pair list out of range");
    const double xab = x0[a] - x0[b], yab = y0[a] - y0[b], zab = z0[a] - z0[b];
    const double rsp = std::sqrt(xab * xab + yab * yab + zab * zab);
    const int ic = (int)(1.0 + (rsp - rlew) / delew);
    ORAC_REQUIRE(ic >= 1, E_ARG, "ERF_CORR: a pair lies below the table's first node");
    icmax = std::max(icmax, ic);
  }
  h.erf_tab = ctx->B.put(pfx + "etab", tab, 4 * (size_t)icmax, ctx->stream, true);
  h.erf_n = icmax;
  h.delew = delew;
  h.rlew = rlew;
}

extern "C" int orac_hip_ferrf(orac_hip_ctx *ctx, const int32_t *ss_index, double alphal, const double *charge,
                              double fudge, const double *x0, const double *y0, const double *z0, int32_t iz,
                              const int32_t *list, int32_t nlist, double *fscnstr_slt, double *fscnstr_slv,
                              double *fpx, double *fpy, double *fpz, int32_t erf_corr, const double *erf_arr_corr,
                              double delew, double rlew, int32_t nbonds_added, int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(ss_index && charge && x0 && y0 && z0 && fscnstr_slt && fscnstr_slv, E_ARG, "null argument");
    ORAC_REQUIRE(mb >= 0 && nlist >= 0 && nbonds_added >= 0 && nbonds_added <= nlist, E_ARG, "bad sizes");
    int n = nlist - nbonds_added;
    ORAC_REQUIRE(n == 0 || list, E_ARG, "null pair list");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    CorrDesc h;
    std::memset(&h, 0, sizeof h);
    if (erf_corr && n > 0) erf_table_for(ctx, h, "c_", erf_arr_corr, delew, rlew, list, n, mb, x0, y0, z0);
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
    const double pi = 4.0 * std::atan(1.0);
    const double erfcst = erfc_as(kcut / (2.0 * alphal));
    for (int n = 0; n < 2; ++n) s[n] = -s[n] * alphal / std::sqrt(pi) * (1.0 - erfcst);
    *self_slt = s[0];
    *self_slv = s[1];
  });
}

extern "C" int orac_hip_int_corr_erf_spline(double rmin, double rmax, int32_t nbin, double alpha, double kcut,
                                            double *erf_arr_corr, double *tau) {
  return guarded([&] {
    ORAC_REQUIRE(nbin >= 1 && erf_arr_corr && tau && rmax > rmin, E_ARG, "bad table arguments");
    int_corr_erf_spline(rmin, rmax, nbin, alpha, kcut, erf_arr_corr, tau);
  });
}

extern "C" int orac_hip_erf_corr_cutoff(const double *oc, double *delew, double *rkcut, int32_t nfft1, int32_t nfft2,
                                        int32_t nfft3) {
  return guarded([&] {
    (void)delew;
    ORAC_REQUIRE(oc && rkcut, E_ARG, "null argument");
    *rkcut = erf_corr_cutoff(oc, nfft1, nfft2, nfft3);
  });
}

extern "C" int orac_hip_ferrf_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, double alphal,
                                      const double *charge, const double *charge0, double fudge, const double *x0,
                                      const double *y0, const double *z0, int32_t iz, const int32_t *list,
                                      int32_t nlist, double *fscnstr_slt, double *fscnstr_slv, double *fpx,
                                      double *fpy, double *fpz, int32_t erf_corr, int32_t nbonds_added, double *dwrk,
                                      int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(!erf_corr, E_UNSUPPORTED, "ferrf_alchemy: the ERF_CORR spline correction is not supported");
    ORAC_REQUIRE(ss_index && charge && charge0 && x0 && y0 && z0 && fscnstr_slt && fscnstr_slv && dwrk, E_ARG,
                 "null argument");
    ORAC_REQUIRE(mb >= 0 && nlist >= 0 && nbonds_added >= 0 && nbonds_added <= nlist, E_ARG, "bad sizes");
    int n = nlist - nbonds_added;
    ORAC_REQUIRE(n == 0 || list, E_ARG, "null pair list");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    CorrDesc h;
    std::memset(&h, 0, sizeof h);
    h.x = B.put("ca_x", x0, mb, st, false);
    h.y = B.put("ca_y", y0, mb, st, false);
    h.z = B.put("ca_z", z0, mb, st, false);
    h.q = B.put("ca_q", charge, mb, st, false);
    h.q0 = B.put("ca_q0", charge0, mb, st, false);
    h.ss = B.put("ca_ss", ss_index, mb, st, true);
    h.alpha = alphal;
    h.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
    h.fudge = fudge;
    h.iz = iz == 1 ? 1 : 0;
    h.kind = 0;
    double e[4];
    run_corr_pairs(ctx, h, list, n, mb, "ca_", e, fpx, fpy, fpz);
    *fscnstr_slt = e[0];
    *fscnstr_slv = e[1];
    *dwrk = e[0] + e[1] - e[2];  // ferrf_alchemy.f: fsrtal(1) + fsrtal(2) - en0
  });
}

extern "C" int orac_hip_ferrf_alchemy_add(orac_hip_ctx *ctx, const int32_t *ss_index, double alphal,
                                          const double *charge, const double *lambdaq, const double *lambdaq0,
                                          const double *x0, const double *y0, const double *z0, const int32_t *list,
                                          int32_t nlist, double *fscnstr_slt, double *fscnstr_slv, double *fpx,
                                          double *fpy, double *fpz, int32_t erf_corr, const double *erf_arr_corr,
                                          double delew, double rlew, double *dwrk, int32_t nato, int32_t intramol) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(ss_index && charge && lambdaq && lambdaq0 && x0 && y0 && z0 && fscnstr_slt && fscnstr_slv && dwrk &&
                     fpx && fpy && fpz,
                 E_ARG, "null argument");
    ORAC_REQUIRE(nato >= 0 && nlist >= 0, E_ARG, "bad sizes");
    ORAC_REQUIRE(nlist == 0 || list, E_ARG, "null pair list");
    for (int p = 0; p < nlist; ++p)
      ORAC_REQUIRE(list[3 * p + 2] == 0 || list[3 * p + 2] == 1, E_ARG, "ferrf_alchemy_add: facQ must be 0 or 1");
    for (int a = 0; a < nato; ++a)
      ORAC_REQUIRE(ss_index[a] == 1 || ss_index[a] == 2, E_ARG, "ss_index must be 1 or 2");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    CorrDesc h;
    std::memset(&h, 0, sizeof h);
    if (erf_corr && nlist > 0)
      erf_table_for(ctx, h, "cq_", erf_arr_corr, delew, rlew, list, nlist, nato, x0, y0, z0, 3);
    std::vector<int> facq(std::max(1, nlist));
    for (int p = 0; p < nlist; ++p) facq[p] = list[3 * p + 2];
    h.x = B.put("cq_x", x0, nato, st, false);
    h.y = B.put("cq_y", y0, nato, st, false);
    h.z = B.put("cq_z", z0, nato, st, false);
    h.qa = B.put("cq_q", charge, nato, st, true);
    h.q = h.qa;
    h.lq = B.put("cq_lq", lambdaq, nato, st, false);
    h.lq0 = B.put("cq_lq0", lambdaq0, nato, st, false);
    h.ss = B.put("cq_ss", ss_index, nato, st, true);
    h.facq = B.put_vec("cq_facq", facq, st, true);
    h.alpha = alphal;
    h.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
    h.iz = 1;
    h.kind = 3;
    h.ncol = 8;
    h.intramol = intramol ? 1 : 0;
    double e[8];
    run_corr_pairs(ctx, h, list, nlist, nato, "cq_", e, fpx, fpy, fpz, 3);
    *fscnstr_slt = e[4];
    *fscnstr_slv = e[5];
    *dwrk = e[4] + e[5] - e[6];  // ferrf_alchemy_add.f:201-203: fsrtal(1) + fsrtal(2) - en0
  });
}

extern "C" int orac_hip_cself_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, int32_t nato, double alphal,
                                      double kcut, const double *charge, const double *charge0, double *self_slt,
                                      double *dwrk) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(ss_index && charge && charge0 && self_slt && dwrk && nato >= 0, E_ARG, "bad argument");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    int *ss = B.put("csa_ss", ss_index, nato, st, false);
    double *q = B.put("csa_q", charge, nato, st, false);
    double *q0 = B.put("csa_q0", charge0, nato, st, false);
    double *dout = B.get<double>("csa_out", 4);
    launch_cself(nato, q, ss, dout, st);
    launch_cself(nato, q0, ss, dout + 2, st);
    double s[4];
    ORAC_HIP(hipMemcpyAsync(s, dout, sizeof s, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    // cself_alchemy.f:36-56: solute atoms only
    const double pi = 4.0 * std::atan(1.0);
    const double erfcst = erfc_as(kcut / (2.0 * alphal));
    const double fs = -s[0] * alphal / std::sqrt(pi) * (1.0 - erfcst);
    const double en0 = -s[2] * alphal / std::sqrt(pi) * (1.0 - erfcst);
    *self_slt = fs;
    *dwrk = fs - en0;
  });
}

// True when `nb` bytes at `host` differ from the copy cached under `name`
// (and then the cache holds the new bytes).
static bool bytes_changed(BufSet &B, const std::string &name, const void *host, size_t nb) {
  auto &c = B.cache[name];
  if (c.size() == nb && (nb == 0 || std::memcmp(c.data(), host, nb) == 0)) return false;
  c.assign(reinterpret_cast<const char *>(host), reinterpret_cast<const char *>(host) + nb);
  return true;
}

// Bonded terms of `arity` atoms each (lst(arity, n), 1-based) on the
// coordinates as given: per-term force components on the device, gathered per
// atom in term order and added to fp; e = energies by the class of each term's
// first atom.  pots: pota, potb (and potc..potg for bends).  The topology, the
// parameters and the per-atom gather index are rebuilt and uploaded only when
// their bytes change (these routines run at the innermost r-RESPA level with
// the same topology every call); the coordinates go up on every call.
// Returns 0, 1 (a singular torsion) or 2 (an unsupported multiplicity).
static int bonded_eval(orac_hip_ctx *ctx, const char *name, int arity, const int32_t *ss_index, const int32_t *lst,
                       int32_t n, const double *xp0, const double *yp0, const double *zp0,
                       const double *const *pots, int npots, double *fpx, double *fpy, double *fpz, int32_t mb,
                       double *e, bool improper = false, const int32_t *remf = nullptr, double remf2 = 1.0,
                       bool rem_variant = false) {
  ORAC_REQUIRE(ss_index && lst && xp0 && yp0 && zp0 && fpx && fpy && fpz && mb > 0 && n > 0, E_ARG, "null argument");
  for (int k = 0; k < npots; ++k) ORAC_REQUIRE(pots[k], E_ARG, "null argument");
  hipStream_t st = ctx->stream;
  BufSet &B = ctx->B;
  const std::string pfx = std::string("t_") + name + "_";
  std::vector<char> key((size_t)n * arity * sizeof(int32_t) + sizeof(int32_t));
  std::memcpy(key.data(), lst, (size_t)n * arity * sizeof(int32_t));
  std::memcpy(key.data() + key.size() - sizeof(int32_t), &mb, sizeof(int32_t));
  if (bytes_changed(B, pfx + "#topo", key.data(), key.size()) || !B.bufs.count(pfx + "gs")) {
    std::vector<int4> at(n);
    std::vector<int> cnt(mb + 1, 0);
    for (int i = 0; i < n; ++i) {
      int l[4] = {0, 0, 0, 0};
      for (int k = 0; k < arity; ++k) {
        l[k] = lst[(size_t)arity * i + k] - 1;
        if (l[k] < 0 || l[k] >= mb) {
          B.cache.erase(pfx + "#topo");
          ORAC_REQUIRE(false, E_ARG, std::string(name) + ": atom index out of range");
        }
        ++cnt[l[k] + 1];
      }
      at[i] = make_int4(l[0], l[1], l[2], l[3]);
    }
    for (int a = 0; a < mb; ++a) cnt[a + 1] += cnt[a];
    std::vector<int> gs(std::max(1, cnt[mb])), cur(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < arity; ++k) gs[cur[lst[(size_t)arity * i + k] - 1]++] = 4 * i + k;
    B.put_vec(pfx + "at", at, st, false);
    B.put_vec(pfx + "goff", cnt, st, false);
    B.put_vec(pfx + "gs", gs, st, false);
  }
  if (bytes_changed(B, pfx + "#ss", ss_index, (size_t)mb * sizeof(int32_t))) {
    for (int a = 0; a < mb; ++a)
      if (ss_index[a] != 1 && ss_index[a] != 2) {
        B.cache.erase(pfx + "#ss");
        ORAC_REQUIRE(false, E_ARG, "ss_index must be 1 or 2");
      }
    B.put(pfx + "ss", ss_index, mb, st, false);
  }
  static const char *pn[7] = {"pa", "pb", "pc", "pd", "pe", "pf", "pg"};
  const double *dp[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int k = 0; k < npots; ++k) dp[k] = B.put(pfx + pn[k], pots[k], n, st, true);
  TorsDesc h;
  std::memset(&h, 0, sizeof h);
  h.ntors = n;
  h.natoms = mb;
  h.arity = arity;
  h.improper = improper ? 1 : 0;
  h.atoms = B.get<int4>(pfx + "at", n);
  h.pota = dp[0];
  h.potb = dp[1];
  h.potc = dp[2];
  h.potd = dp[3];
  h.pote = dp[4];
  h.potf = dp[5];
  h.potg = dp[6];
  h.x = B.put("t_x", xp0, mb, st, false);
  h.y = B.put("t_y", yp0, mb, st, false);
  h.z = B.put("t_z", zp0, mb, st, false);
  h.ss = B.get<int>(pfx + "ss", mb);
  h.tf = B.get<double>("t_tf", 12 * (size_t)n);
  h.te = B.get<double>("t_te", 2 * (size_t)n);
  h.status = B.get<int>("t_status", 2);
  launch_zero_u64(reinterpret_cast<unsigned long long *>(h.status), 1, st);  // both status words
  h.g_off = B.get<int>(pfx + "goff", mb + 1);
  h.g_slot = B.get<int>(pfx + "gs", 1);
  h.fx = B.get<double>("t_fx", mb);
  h.fy = B.get<double>("t_fy", mb);
  h.fz = B.get<double>("t_fz", mb);
  // REM twins: e[2] returns the flagged terms' energy
  const int ne = remf ? 3 : 2;
  if (remf) {
    std::vector<int> fl(n);
    for (int i = 0; i < n; ++i) fl[i] = remf[i] != 0;
    h.remf = B.put_vec("t_remf", fl, st, false);
    h.remf2 = remf2;
    h.rem_variant = rem_variant ? 1 : 0;
    h.tseg = B.get<double>("t_tseg", n);
  }
  double *dout = B.get<double>("t_out", 3);
  h.out = dout;
  TorsDesc *dd = B.put("t_desc", &h, 1, st, false);
  launch_tors(dd, h, st);
  ORAC_HIP(hipGetLastError());
  std::vector<double> fx(mb), fy(mb), fz(mb);
  int status[2] = {0, 0};
  ORAC_HIP(hipMemcpyAsync(fx.data(), h.fx, mb * sizeof(double), hipMemcpyDeviceToHost, st));
  ORAC_HIP(hipMemcpyAsync(fy.data(), h.fy, mb * sizeof(double), hipMemcpyDeviceToHost, st));
  ORAC_HIP(hipMemcpyAsync(fz.data(), h.fz, mb * sizeof(double), hipMemcpyDeviceToHost, st));
  ORAC_HIP(hipMemcpyAsync(e, dout, ne * sizeof(double), hipMemcpyDeviceToHost, st));
  ORAC_HIP(hipMemcpyAsync(status, h.status, sizeof status, hipMemcpyDeviceToHost, st));
  ORAC_HIP(hipStreamSynchronize(st));
  // an unsupported multiplicity (the reference stops) outranks a singular torsion
  const int code = status[1] ? 2 : (status[0] ? 1 : 0);
  if (code == 0)
    for (int a = 0; a < mb; ++a) {
      fpx[a] = fpx[a] + fx[a];
      fpy[a] = fpy[a] + fy[a];
      fpz[a] = fpz[a] + fz[a];
    }
  return code;
}

extern "C" int orac_hip_fptors(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *ltor, int32_t ltors,
                               const double *xp0, const double *yp0, const double *zp0, const double *pota,
                               const double *potb, double *utors_slt, double *utors_slv, double *fpx, double *fpy,
                               double *fpz, int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(utors_slt && utors_slv, E_ARG, "null output");
    *utors_slt = *utors_slv = 0.0;
    if (ltors == 0) return;
    ORAC_REQUIRE(ltors > 0, E_ARG, "negative count");
    const double *pots[2] = {pota, potb};
    double e[2];
    const int status = bonded_eval(ctx, "fptors", 4, ss_index, ltor, ltors, xp0, yp0, zp0, pots, 2, fpx, fpy, fpz,
                                   mb, e);
    ORAC_REQUIRE(status != 2, E_UNSUPPORTED, "fptors: torsion multiplicity outside 1-8, 12, 14");
    ORAC_REQUIRE(status != 1, E_STATE,
                 "fptors: a torsion with collinear atoms (the reference perturbs it at random; not reproduced)");
    *utors_slt = e[0];
    *utors_slv = e[1];
  });
}

extern "C" int orac_hip_fitors(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *litr, int32_t litor,
                               const double *xp0, const double *yp0, const double *zp0, const double *pota,
                               const double *potb, const double *potc, double *utors_slt, double *utors_slv,
                               double *fpx, double *fpy, double *fpz, int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(utors_slt && utors_slv, E_ARG, "null output");
    *utors_slt = *utors_slv = 0.0;
    if (litor == 0) return;
    ORAC_REQUIRE(litor > 0, E_ARG, "negative count");
    const double *pots[3] = {pota, potb, potc};
    double e[2];
    const int status = bonded_eval(ctx, "fitors", 4, ss_index, litr, litor, xp0, yp0, zp0, pots, 3, fpx, fpy, fpz,
                                   mb, e, true);
    ORAC_REQUIRE(status != 2, E_UNSUPPORTED, "fitors: improper torsion multiplicity other than 1, 2, 3, 4, 6");
    *utors_slt = e[0];
    *utors_slv = e[1];
  });
}

extern "C" int orac_hip_fptors_rem(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *ltor, int32_t ltors,
                                   const double *xp0, const double *yp0, const double *zp0, const double *pota,
                                   const double *potb, double *utors_slt, double *utors_slv, double *fpx, double *fpy,
                                   double *fpz, const int32_t *remltor, double *seg_ene, double rem_factor2,
                                   int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(utors_slt && utors_slv && seg_ene, E_ARG, "null output");
    *utors_slt = *utors_slv = *seg_ene = 0.0;
    if (ltors == 0) return;
    ORAC_REQUIRE(ltors > 0 && remltor, E_ARG, "negative count or null remltor");
    const double *pots[2] = {pota, potb};
    double e[3];
    const int status = bonded_eval(ctx, "fptors", 4, ss_index, ltor, ltors, xp0, yp0, zp0, pots, 2, fpx, fpy, fpz,
                                   mb, e, false, remltor, rem_factor2, true);
    ORAC_REQUIRE(status != 2, E_UNSUPPORTED, "fptors_rem: torsion multiplicity outside 1-8, 12, 14");
    ORAC_REQUIRE(status != 1, E_STATE,
                 "fptors_rem: a torsion with collinear atoms (the reference perturbs it at random; not reproduced)");
    *utors_slt = e[0];
    *utors_slv = e[1];
    *seg_ene = e[2];
  });
}

extern "C" int orac_hip_fitors_rem(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *litr, int32_t litor,
                                   const double *xp0, const double *yp0, const double *zp0, const double *pota,
                                   const double *potb, const double *potc, double *utors_slt, double *utors_slv,
                                   double *fpx, double *fpy, double *fpz, const int32_t *remitor, double *seg_ene,
                                   double rem_factor2, int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(utors_slt && utors_slv && seg_ene, E_ARG, "null output");
    *utors_slt = *utors_slv = *seg_ene = 0.0;
    if (litor == 0) return;
    ORAC_REQUIRE(litor > 0 && remitor, E_ARG, "negative count or null remitor");
    const double *pots[3] = {pota, potb, potc};
    double e[3];
    const int status = bonded_eval(ctx, "fitors", 4, ss_index, litr, litor, xp0, yp0, zp0, pots, 3, fpx, fpy, fpz,
                                   mb, e, true, remitor, rem_factor2);
    ORAC_REQUIRE(status != 2, E_UNSUPPORTED, "fitors_rem: improper torsion multiplicity other than 1, 2, 3, 4, 6");
    *utors_slt = e[0];
    *utors_slv = e[1];
    *seg_ene = e[2];
  });
}

extern "C" int orac_hip_fpbond(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *lbnd, int32_t lbond,
                               const double *xp0, const double *yp0, const double *zp0, const double *pota,
                               const double *potb, double *ubond_slt, double *ubond_slv, double *fpx, double *fpy,
                               double *fpz, int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(ubond_slt && ubond_slv, E_ARG, "null output");
    *ubond_slt = *ubond_slv = 0.0;
    if (lbond == 0) return;
    ORAC_REQUIRE(lbond > 0, E_ARG, "negative count");
    const double *pots[2] = {pota, potb};
    double e[2];
    bonded_eval(ctx, "fpbond", 2, ss_index, lbnd, lbond, xp0, yp0, zp0, pots, 2, fpx, fpy, fpz, mb, e);
    *ubond_slt = e[0];
    *ubond_slv = e[1];
  });
}

extern "C" int orac_hip_fpbend(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *lbndg, int32_t lbend,
                               const double *xp0, const double *yp0, const double *zp0, const double *pota,
                               const double *potb, const double *potc, const double *potd, const double *pote,
                               const double *potf, const double *potg, double *ubend_slt, double *ubend_slv,
                               double *fpx, double *fpy, double *fpz, int32_t mb) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(ubend_slt && ubend_slv, E_ARG, "null output");
    *ubend_slt = *ubend_slv = 0.0;
    if (lbend == 0) return;
    ORAC_REQUIRE(lbend > 0, E_ARG, "negative count");
    const double *pots[7] = {pota, potb, potc, potd, pote, potf, potg};
    double e[2];
    bonded_eval(ctx, "fpbend", 3, ss_index, lbndg, lbend, xp0, yp0, zp0, pots, 7, fpx, fpy, fpz, mb, e);
    *ubend_slt = e[0];
    *ubend_slv = e[1];
  });
}

extern "C" int orac_hip_fnb14(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                              const double *zp0, const double *charge, int32_t nato, const double *ecc12,
                              const double *ecc6, int32_t cut, int32_t ewald, double alphal, const int32_t *int14,
                              int32_t int14p, const int32_t *type14, double fudge, double lj_fudge, double *fpx,
                              double *fpy, double *fpz, double *uconf_slt, double *uconf_slv, double *ucoul_slt,
                              double *ucoul_slv) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(uconf_slt && uconf_slv && ucoul_slt && ucoul_slv, E_ARG, "null output");
    *uconf_slt = *uconf_slv = *ucoul_slt = *ucoul_slv = 0.0;
    if (int14p == 0) return;  // fnb14.f:64
    ORAC_REQUIRE(ss_index && xp0 && yp0 && zp0 && charge && ecc12 && ecc6 && int14 && type14 && nato > 0 &&
                     int14p > 0,
                 E_ARG, "null argument");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    int lmax = 1;
    for (int k = 0; k < int14p; ++k) {
      ORAC_REQUIRE(type14[k] >= 1, E_ARG, "fnb14: type14 entries are 1-based packed indices");
      lmax = std::max(lmax, type14[k]);
    }
    for (int a = 0; a < nato; ++a)
      ORAC_REQUIRE(ss_index[a] == 1 || ss_index[a] == 2, E_ARG, "ss_index must be 1 or 2");
    CorrDesc h;
    std::memset(&h, 0, sizeof h);
    h.xp = B.put("n_x", xp0, nato, st, false);
    h.yp = B.put("n_y", yp0, nato, st, false);
    h.zp = B.put("n_z", zp0, nato, st, false);
    h.q = B.put("n_q", charge, nato, st, true);
    h.ss = B.put("n_ss", ss_index, nato, st, true);
    h.pair_type = B.put("n_t14", type14, int14p, st, true);
    h.ecc12 = B.put("n_e12", ecc12, lmax, st, true);
    h.ecc6 = B.put("n_e6", ecc6, lmax, st, true);
    h.alpha = alphal;
    h.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
    h.fudge = fudge;
    h.lj_fudge = lj_fudge;
    h.iz = 1;
    h.kind = 2;
    h.ewald = (ewald && cut) ? 1 : 0;  // fnb14.f:67: plain Coulomb unless ewald and cut
    double e[4];
    run_corr_pairs(ctx, h, int14, int14p, nato, "n_", e, fpx, fpy, fpz);
    *ucoul_slt = e[0];
    *ucoul_slv = e[1];
    *uconf_slt = e[2];
    *uconf_slv = e[3];
  });
}

extern "C" int orac_hip_fnb14_rem(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                                  const double *zp0, const double *charge, int32_t nato, const double *ecc12,
                                  const double *ecc6, int32_t cut, int32_t ewald, double alphal, const int32_t *int14,
                                  int32_t int14p, const int32_t *type14, double fudge, double lj_fudge, double *fpx,
                                  double *fpy, double *fpz, double *uconf_slt, double *uconf_slv, double *ucoul_slt,
                                  double *ucoul_slv, const int32_t *remint14, double *seg_ene, double rem_factor2) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(uconf_slt && uconf_slv && ucoul_slt && ucoul_slv && seg_ene, E_ARG, "null output");
    *uconf_slt = *uconf_slv = *ucoul_slt = *ucoul_slv = 0.0;
    if (int14p == 0) return;  // fnb14_rem.f:100, before seg_ene is zeroed (it is left as it was)
    *seg_ene = 0.0;
    ORAC_REQUIRE(ss_index && xp0 && yp0 && zp0 && charge && ecc12 && ecc6 && int14 && type14 && remint14 &&
                     nato > 0 && int14p > 0,
                 E_ARG, "null argument");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    int lmax = 1;
    for (int k = 0; k < int14p; ++k) {
      ORAC_REQUIRE(type14[k] >= 1, E_ARG, "fnb14_rem: type14 entries are 1-based packed indices");
      lmax = std::max(lmax, type14[k]);
    }
    for (int a = 0; a < nato; ++a)
      ORAC_REQUIRE(ss_index[a] == 1 || ss_index[a] == 2, E_ARG, "ss_index must be 1 or 2");
    std::vector<int> fl(int14p);
    for (int k = 0; k < int14p; ++k) fl[k] = remint14[k] != 0;
    CorrDesc h;
    std::memset(&h, 0, sizeof h);
    h.xp = B.put("n_x", xp0, nato, st, false);
    h.yp = B.put("n_y", yp0, nato, st, false);
    h.zp = B.put("n_z", zp0, nato, st, false);
    h.q = B.put("n_q", charge, nato, st, true);
    h.ss = B.put("n_ss", ss_index, nato, st, true);
    h.pair_type = B.put("n_t14", type14, int14p, st, true);
    h.ecc12 = B.put("n_e12", ecc12, lmax, st, true);
    h.ecc6 = B.put("n_e6", ecc6, lmax, st, true);
    h.remf = B.put_vec("n_remf", fl, st, false);
    h.remf2 = rem_factor2;
    h.ncol = 5;
    h.alpha = alphal;
    h.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
    h.fudge = fudge;
    h.lj_fudge = lj_fudge;
    h.iz = 1;
    h.kind = 2;
    h.ewald = (ewald && cut) ? 1 : 0;
    double e[5];
    run_corr_pairs(ctx, h, int14, int14p, nato, "nr_", e, fpx, fpy, fpz);
    *ucoul_slt = e[0];
    *ucoul_slv = e[1];
    *uconf_slt = e[2];
    *uconf_slv = e[3];
    *seg_ene = e[4];
  });
}

extern "C" int orac_hip_fnb14_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0,
                                      const double *yp0, const double *zp0, const double *charge, int32_t nato,
                                      const double *ecc12, const double *ecc6, int32_t cut, int32_t ewald,
                                      double alphal, const int32_t *int14, int32_t int14p, const int32_t *type14,
                                      double fudge, double lj_fudge, double *fpx, double *fpy, double *fpz,
                                      double *uconf_slt, double *uconf_slv, double *ucoul_slt, double *ucoul_slv,
                                      const double *lambda, const double *lambdaq, const double *lambda0,
                                      const double *lambdaq0, double *dwrk) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(uconf_slt && uconf_slv && ucoul_slt && ucoul_slv && dwrk, E_ARG, "null output");
    *uconf_slt = *uconf_slv = *ucoul_slt = *ucoul_slv = 0.0;
    *dwrk = 0.0;  // fnb14_alchemy.f:117 (serial build; mts_intra_n1.f:61 zeroes it in the OpenMP one)
    if (int14p == 0) return;
    ORAC_REQUIRE(ss_index && xp0 && yp0 && zp0 && charge && ecc12 && ecc6 && int14 && type14 && lambda && lambdaq &&
                     lambda0 && lambdaq0 && fpx && fpy && fpz && nato > 0 && int14p > 0,
                 E_ARG, "null argument");
    hipStream_t st = ctx->stream;
    BufSet &B = ctx->B;
    int lmax = 1;
    for (int k = 0; k < int14p; ++k) {
      ORAC_REQUIRE(type14[k] >= 1, E_ARG, "fnb14_alchemy: type14 entries are 1-based packed indices");
      lmax = std::max(lmax, type14[k]);
    }
    for (int a = 0; a < nato; ++a)
      ORAC_REQUIRE(ss_index[a] == 1 || ss_index[a] == 2, E_ARG, "ss_index must be 1 or 2");
    CorrDesc h;
    std::memset(&h, 0, sizeof h);
    h.xp = B.put("na_x", xp0, nato, st, false);
    h.yp = B.put("na_y", yp0, nato, st, false);
    h.zp = B.put("na_z", zp0, nato, st, false);
    h.q = B.put("na_q", charge, nato, st, true);
    h.ss = B.put("na_ss", ss_index, nato, st, true);
    h.pair_type = B.put("na_t14", type14, int14p, st, true);
    h.ecc12 = B.put("na_e12", ecc12, lmax, st, true);
    h.ecc6 = B.put("na_e6", ecc6, lmax, st, true);
    h.lam = B.put("na_l", lambda, nato, st, false);
    h.lq = B.put("na_lq", lambdaq, nato, st, false);
    h.lam0 = B.put("na_l0", lambda0, nato, st, false);
    h.lq0 = B.put("na_lq0", lambdaq0, nato, st, false);
    h.alpha = alphal;
    h.twrtpi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
    h.fudge = fudge;
    h.lj_fudge = lj_fudge;
    h.iz = 1;
    h.kind = 4;
    h.ncol = 8;
    h.ewald = (ewald && cut) ? 1 : 0;  // fnb14_alchemy.f:119: the erfc branch needs ewald and cut
    double e[8];
    run_corr_pairs(ctx, h, int14, int14p, nato, "na_", e, fpx, fpy, fpz);
    *ucoul_slt = e[0];
    *ucoul_slv = e[1];
    *uconf_slt = e[2];
    *uconf_slv = e[3];
    double dw = 0.0;  // fnb14_alchemy.f:349-351, class by class
    for (int k = 0; k < 2; ++k) dw = dw + e[2 + k] + e[k] - (e[6 + k] + e[4 + k]);
    *dwrk = dw;
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


extern "C" int orac_hip_fnbgrp_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0,
                                       const double *yp0, const double *zp0, const double *charge, const int32_t *type,
                                       const double *ecc12, const double *ecc6, int32_t ewald, double alphal,
                                       const int32_t *ingrp, int32_t ingrpp, double *fpx, double *fpy, double *fpz,
                                       double *uconf_slt, double *uconf_slv, double *ucoul_slt, double *ucoul_slv,
                                       const double *lambdaq, int32_t nato) {
  // fnbgrp_alchemy.f:1-220 computes alphal0 from lambdaq (:152-157) and never
  // reads it: its arithmetic is fnbgrp.f's, term for term
  (void)lambdaq;
  return orac_hip_fnbgrp(ctx, ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, ewald, alphal, ingrp, ingrpp, fpx,
                         fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, nato);
}
