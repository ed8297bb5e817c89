// engine.cpp -- device-resident systems and batched evaluation.
//
// A system keeps its topology, coordinates, neighbour list, grids and
// results in HBM.  A batch is an ordered set of systems evaluated together:
// each stage of the force evaluation is one kernel launch for the whole batch
// (blockIdx.y selects the system), so 32 replicas of a 24.5k-atom system fill
// the chip the way one system cannot.  A batch owns copies of its systems'
// descriptors (with the batch's tuning applied) on the host and the device,
// optional captured HIP graphs of its step, and stage timing events.
//
// Systems are independent in a batch: every reduction is per system and in
// fixed order, so a system's forces are the same alone and batched, and its
// energies differ at most by the FFT tiling of the batch's tuning.
#include <chrono>
#include <cmath>

#include "host.hpp"

namespace orac {

enum Stage { ST_BEGIN, ST_CENTRES, ST_LIST, ST_PAIRS, ST_PME, ST_END, ST_KBEG, ST_KEND, ST_P0, ST_P1, ST_P2, ST_P3,
             NSTAGE };
constexpr int NSTAGE_MS = 5;  // centres, list, pairs, pme, corrections; then the pair kernel alone

struct System {
  BufSet B;
  int natoms = 0, ngrp = 0, nmol = 0, ntypes = 0, max_group = 1;
  int precision = 0;
  double co[9], oc[9], volume = 0, alpha = 0;
  double rcut = 0, rtol = 0, rnei = 0;
  PmePlan plan;
  PmeDesc pme;
  PairDesc pair;
  ListDesc list;
  CorrDesc corr;
  size_t list_cap = 0;
  int slab = 0;           // neighbour-list entries per owner (slab layout)
  double *xp = nullptr, *yp = nullptr, *zp = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;
  double *out = nullptr;  // device: pme[20], pairs[16] at 20, corrections[4] at 36, lambda0 pme[20] at 48
  double self_e[2] = {0, 0};
  std::vector<double> q_host;      // charges as created
  std::vector<int32_t> ss_host;    // ss_index
  // alchemy (orac_hip_system_set_alchemy): soft-core pairs on the created
  // charges, reciprocal sum and corrections on the lambda charges, and an
  // energy-only grid on the lambda0 charges
  bool alch = false;
  std::vector<int32_t> ss_alch;
  bool intramol = false, ljintra = false, shifted = false;
  double qsoft = 0.5, ljsoft = 0.5;  // (qshift, ljshift) when shifted
  double lam[8] = {1, 0, 1, 0, 1, 0, 1, 0};
  double self_dw = 0.0;
  double *qc = nullptr, *q0 = nullptr, *afac = nullptr;
  // per-step reciprocal corrections as created (ferrf pairs, 0-based) and, for
  // an alchemical system, the intra-alchemical pairs of ferrf_alchemy_add
  std::vector<int> corr_ia, corr_ib;
  std::vector<double> corr_fud;
  int nqq = 0;
  double *lq = nullptr, *lq0 = nullptr;  // per-atom lambdaq, lambdaq0 (device)
  PmeDesc pme0;
  bool have_list = false;
  int list_fmt = -1;      // cell_order tuning the current list was built with
  int solo = 0;           // id of this system's own one-system batch (0: none yet)
  int gen = 0;            // bumped when a device buffer behind the descriptors moves
  // molecular dynamics (set up by orac_hip_md_setup)
  bool md_on = false;
  bool md_primed = false;  // slow and fast forces at the current positions are in place
  MdDesc md;
  MdDesc *md_d = nullptr;  // device copy for per-system kernels
  // two nonbonded shells (orac_hip_system_set_shells, ORAC's m and outer
  // shells, mts_forces.f:66-93): `pair` is the outer shell (rcutm .. rcut,
  // switched in over rtolm), pair_m the inner one (0 .. rcutm, switched out
  // over rtolm) with forces fm* and energies out_m of its own
  double rcutm = 0.0, rtolm = 0.0;
  PairDesc pair_m;
  double *fmx = nullptr, *fmy = nullptr, *fmz = nullptr, *out_m = nullptr;
};

// Neighbour-side force slots of the half-list form: at least `need`, in three
// FP64 arrays (grown, never shrunk; the descriptors follow through gen).
static void size_half_slots(System &S, long long need) {
  if (need <= S.list.pcap) return;
  const size_t cap = (size_t)need;
  S.pair.jpx = S.B.get<double>("jpx", cap);
  S.pair.jpy = S.B.get<double>("jpy", cap);
  S.pair.jpz = S.B.get<double>("jpz", cap);
  S.list.pcap = need;
  S.pair.pcap = need;
  S.gen++;
}

struct Batch {
  std::vector<int> ids;
  std::vector<System *> sys;
  std::vector<int> gens;  // system generations the descriptor copies were made from
  int precision = 0;
  Tuning tu;
  std::vector<std::pair<Tuning, double>> tune_log;  // every candidate the last autotune timed, ms per step
  BufSet B;
  std::vector<PmeDesc> hp;
  std::vector<PairDesc> hq;
  std::vector<ListDesc> hl;
  std::vector<CorrDesc> hc;
  std::vector<PairDesc> hqm;  // inner-shell descriptors (every system has two shells)
  PairDesc *dqm = nullptr;
  bool shells = false;
  PmeDesc *dp = nullptr;
  PairDesc *dq = nullptr;
  ListDesc *dl = nullptr;
  CorrDesc *dc = nullptr;
  bool any_corr = false;
  hipEvent_t fork = nullptr, join = nullptr;  // reciprocal sum on the side stream
  hipGraph_t graph[2] = {nullptr, nullptr};  // [rebuild]
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  bool profile = false;
  hipEvent_t ev[NSTAGE] = {};
  // molecular dynamics
  std::vector<MdDesc> hm;
  MdDesc *dm = nullptr;
  std::vector<ListDesc> hl_md;  // list descriptors driven by each system's step counter
  ListDesc *dl_md = nullptr;
  long md_gen = -1;             // engine md generation the descriptors were built from
  int md_every = 0;             // list update period the md descriptors carry
  double md_dt = 0.0;
  int md_ninner = 0, md_nmid = 0;
  hipGraph_t mgraph = nullptr;  // one outer step of every system
  hipGraphExec_t mexec = nullptr;
  void drop_md_graphs() {
    if (mexec) (void)hipGraphExecDestroy(mexec);
    if (mgraph) (void)hipGraphDestroy(mgraph);
    mexec = nullptr;
    mgraph = nullptr;
  }
  ~Batch() {
    drop_md_graphs();
    for (int k = 0; k < 2; ++k) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      if (graph[k]) (void)hipGraphDestroy(graph[k]);
    }
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
  }
};

struct Engine {
  std::map<int, std::unique_ptr<System>> systems;
  std::map<int, std::unique_ptr<Batch>> batches;
  int next_sys = 1, next_batch = 1;
  long md_gen = 0;  // bumped by every md setup or state change
};

void engine_destroy(Engine *e) { delete e; }

static Engine &engine(orac_hip_ctx *ctx) {
  ORAC_REQUIRE(ctx, E_ARG, "null context");
  if (!ctx->engine) ctx->engine = new Engine();
  return *ctx->engine;
}

static System &get_sys(orac_hip_ctx *ctx, int32_t id) {
  Engine &E = engine(ctx);
  auto it = E.systems.find(id);
  ORAC_REQUIRE(it != E.systems.end(), E_ARG, "no such system");
  return *it->second;
}

static Batch &get_batch(orac_hip_ctx *ctx, int32_t id) {
  Engine &E = engine(ctx);
  auto it = E.batches.find(id);
  ORAC_REQUIRE(it != E.batches.end(), E_ARG, "no such batch");
  return *it->second;
}

// ---------------------------------------------------------------- system build

// The system's correction list: its ferrf pairs (kind 0) and, when given, the
// intra-alchemical pairs qq(3, nqq) of ferrf_alchemy_add (kind 3, 1-based with
// facQ), in that order, with the per-atom gather index in pair order.
static void corr_upload(System &S, const int32_t *qq, int nqq, hipStream_t st) {
  BufSet &B = S.B;
  CorrDesc &h = S.corr;
  const int n = S.natoms, nc = (int)S.corr_ia.size(), np = nc + nqq;
  std::vector<int> ia(std::max(1, np)), ib(std::max(1, np)), kind(std::max(1, np), 0), facq(std::max(1, np), 0);
  std::vector<double> fud(std::max(1, np), 1.0);
  for (int k = 0; k < nc; ++k) {
    ia[k] = S.corr_ia[k];
    ib[k] = S.corr_ib[k];
    fud[k] = S.corr_fud[k];
  }
  for (int k = 0; k < nqq; ++k) {
    ia[nc + k] = qq[3 * k] - 1;
    ib[nc + k] = qq[3 * k + 1] - 1;
    facq[nc + k] = qq[3 * k + 2];
    kind[nc + k] = 3;
    ORAC_REQUIRE(ia[nc + k] >= 0 && ia[nc + k] < n && ib[nc + k] >= 0 && ib[nc + k] < n, E_ARG,
                 "intra-alchemical pair out of range");
    ORAC_REQUIRE(facq[nc + k] == 0 || facq[nc + k] == 1, E_ARG, "facQ must be 0 or 1");
  }
  std::vector<int> cnt(n + 1, 0);
  for (int k = 0; k < np; ++k) {
    ++cnt[ia[k] + 1];
    ++cnt[ib[k] + 1];
  }
  for (int a = 0; a < n; ++a) cnt[a + 1] += cnt[a];
  std::vector<int> gp(std::max(1, cnt[n])), cur(cnt.begin(), cnt.end() - 1);
  for (int k = 0; k < np; ++k) {
    gp[cur[ia[k]]++] = k + 1;
    gp[cur[ib[k]]++] = -(k + 1);
  }
  const int ncol = nqq > 0 ? 8 : 4;
  h.npairs = np;
  h.natoms = n;
  h.ncol = ncol;
  h.ia = B.put_vec("cia", ia, st, false);
  h.ib = B.put_vec("cib", ib, st, false);
  h.fudge_p = B.put_vec("cfud", fud, st, false);
  h.pkind = nqq > 0 ? B.put_vec("ckind", kind, st, false) : nullptr;
  h.facq = nqq > 0 ? B.put_vec("cfacq", facq, st, false) : nullptr;
  h.g_off = B.put_vec("cgoff", cnt, st, false);
  h.g_pair = B.put_vec("cgp", gp, st, false);
  h.pf = B.get<double>("cpf", 3 * (size_t)std::max(1, np));
  h.pe = B.get<double>("cpe", ncol * (size_t)std::max(1, np));
  S.nqq = nqq;
}

static void system_build(orac_hip_ctx *ctx, System &S, const orac_hip_system_desc &d) {
  hipStream_t st = ctx->stream;
  BufSet &B = S.B;
  const int n = d.natoms, ng = d.ngrp, ma = d.ntypes;
  ORAC_REQUIRE(n > 0 && ng > 0 && ma > 0 && d.nmol > 0, E_ARG, "system sizes must be positive");
  // device kernels address atom and group records with 32-bit byte offsets
  ORAC_REQUIRE(n < (1 << 26), E_UNSUPPORTED, "more than 2^26 atoms in one system");
  ORAC_REQUIRE(d.charge && d.mass && d.nbtype && d.type_table && d.ecc12 && d.ecc6 && d.grppt && d.atomp &&
                   d.ss_index && d.mapnl && d.co,
               E_ARG, "null array in system description");
  ORAC_REQUIRE(d.precision == 0 || d.precision == 1, E_ARG, "precision must be 0 or 1");
  ORAC_REQUIRE(d.ncorr >= 0 && (d.ncorr == 0 || d.corr_pairs), E_ARG, "bad correction list");
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
  ORAC_REQUIRE(S.max_group <= 64, E_UNSUPPORTED, "groups larger than 64 atoms are not supported");
  std::vector<int> mfirst(d.nmol, -1), mlast(d.nmol, -1);
  for (int a = 0; a < n; ++a) {
    ORAC_REQUIRE(agroup[a] >= 0, E_ARG, "an atom belongs to no group");
    ORAC_REQUIRE(d.ss_index[a] == 1 || d.ss_index[a] == 2, E_ARG, "ss_index must be 1 (solute) or 2 (solvent)");
    tp[a] = d.nbtype[a] - 1;
    ap[a] = d.atomp[a] - 1;
    ORAC_REQUIRE(tp[a] >= 0 && tp[a] < ma && ap[a] >= 0 && ap[a] < d.nmol, E_ARG, "nbtype or atomp out of range");
    if (mfirst[ap[a]] < 0) mfirst[ap[a]] = a;
    ORAC_REQUIRE(mlast[ap[a]] < 0 || mlast[ap[a]] == a - 1, E_ARG, "molecules must be contiguous atom ranges");
    mlast[ap[a]] = a;
  }
  for (int m = 0; m < d.nmol; ++m) ORAC_REQUIRE(mfirst[m] >= 0, E_ARG, "a molecule has no atoms");
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
  S.out = B.get<double>("out", 72);
  launch_zero_f64(S.out, 72, st);
  S.q_host.assign(d.charge, d.charge + n);
  S.ss_host.assign(d.ss_index, d.ss_index + n);

  // ---- direct space
  PairDesc &p = S.pair;
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
  p.nmol = d.nmol;
  p.mfirst = B.put_vec("mfirst", mfirst, st, false);
  p.mlast = B.put_vec("mlast", mlast, st, false);
  {
    std::vector<int> big;
    for (int m = 0; m < d.nmol; ++m)
      if (mlast[m] - mfirst[m] + 1 > MOL_SMALL) big.push_back(m);
    p.nbig = (int)big.size();
    if (big.empty()) big.push_back(0);
    p.mbig = B.put_vec("mbig", big, st, false);
  }
  p.gfirst = B.put_vec("gfirst", gfirst, st, false);
  p.glast = B.put_vec("glast", glast, st, false);
  p.gss = B.put_vec("gss", gss, st, false);
  p.xpg = xpg;
  p.ypg = ypg;
  p.zpg = zpg;
  p.ntypes = ma;
  const size_t npk = (size_t)ma * (ma + 1) / 2;
  if (ma <= LJ_DENSE_MAX) {
    // the indices checked here, the table gathered on the device from the
    // packed arrays (staged in the context's buffers, reused by every system)
    const size_t n2 = (size_t)ma * ma;
    int lo = 1, hi = 1;
    if (n2 > 0) lo = hi = d.type_table[0];
    for (size_t k = 0; k < n2; ++k) {
      lo = std::min(lo, (int)d.type_table[k]);
      hi = std::max(hi, (int)d.type_table[k]);
    }
    ORAC_REQUIRE(lo >= 1 && (size_t)hi <= npk, E_ARG, "type table index out of range");
    int *dtt = ctx->B.put("ljs_tt", d.type_table, n2, st, false);
    double *d12 = ctx->B.put("ljs_e12", d.ecc12, npk, st, false);
    double *d6 = ctx->B.put("ljs_e6", d.ecc6, npk, st, false);
    double2 *dense = B.get<double2>("ljd", n2);
    float2 *dense32 = d.precision == 1 ? B.get<float2>("ljd32", n2) : nullptr;
    launch_lj_dense(ma, dtt, d12, d6, dense, dense32, st);
    p.lj_dense = dense;
    if (dense32) p.lj_dense32 = dense32;
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
  p.split = pair_split_for(ng);
  p.nblk = (ng + PAIR_WAVES / p.split - 1) / (PAIR_WAVES / p.split);
  p.part = B.get<double>("ppart", (size_t)p.nblk * 16);
  p.sgp = B.get<double4>("sgp", ng);
  p.sgi = B.get<int4>("sgi", ng);
  p.sap = B.get<double4>("sap", n);
  if (d.precision == 1) p.sapf = B.get<float4>("sapf", n);
  else p.sapc = B.get<double4>("sapc", n);
  p.sai = B.get<int2>("sai", n);
  p.out = S.out + 20;
  p.max_group = S.max_group;

  // ---- linked cells and the full list
  ListDesc &L = S.list;
  std::memset(&L, 0, sizeof L);
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
    L.cg = make_cellgrid(B, "", ncx, ncy, ncz, nind.data(), ii.data(), jj.data(), kk.data(), rl, d.co, 0, st);
    L.cg.cpos = B.get<double4>("cpos", ng);
    const int ncell = ncx * ncy * ncz;
    const double dens = ng / S.volume;
    const double nb = 4.0 / 3.0 * M_PI * rl * rl * rl * dens;
    // slab layout: room for twice the mean neighbour count per owner until the
    // first coordinates give the real largest count (size_list)
    S.slab = std::min(ng, ((int)(2.0 * nb) + 64 + 31) / 32 * 32);
    S.list_cap = (size_t)ng * (size_t)S.slab;
    L.ngrp = ng;
    L.xpg = xpg;
    L.ypg = ypg;
    L.zpg = zpg;
    L.cell_of = B.get<int>("cellof", ng);
    L.cell_cnt = B.get<int>("ccnt", ncell + 1);
    L.cell_start = B.get<int>("cstart", ncell + 1);
    L.cell_cursor = B.get<int>("ccur", ncell + 1);
    L.cell_items = B.get<int>("citems", ng);
    L.cell_fill = B.get<int>("cfill", ng);
    L.count = B.get<int>("lcount", ng + 1);
    L.off = B.get<int>("nboff", ng + 1);
    L.out = B.get<int>("nbidx", S.list_cap);
    L.cap = (long long)S.list_cap;
    L.total = B.get<int>("ltotal", 4);
    L.mode = 0;
    L.slab = S.slab;
    std::vector<int> gsize(ng);
    for (int g = 0; g < ng; ++g) gsize[g] = glast[g] - gfirst[g] + 1;
    L.gsize = B.put_vec("gsize", gsize, st, false);
    L.soff = B.get<int>("soff", ng);
    launch_zero_u64(reinterpret_cast<unsigned long long *>(L.total), 2, st);
  }
  p.nb_off = L.off;
  p.nb_idx = L.out;
  p.nb_cnt = L.count;
  p.nb_slab = L.slab;
  p.perm = L.cell_items;  // cell order of the last list build
  p.soff = L.soff;
  // half-list form: groups of at most OWN_REG atoms, one wave per owner, non-alchemical
  // (set_alchemy switches back)
  if (use_owner_kernel() && S.max_group <= OWN_REG && p.split == 1) {
    L.split = 1;
    L.rout = B.get<int>("nbrout", S.list_cap);
    L.mirror = B.get<int>("nbmirror", S.list_cap);
    L.rcount = B.get<int>("nbrcount", ng);
    L.rbase = B.get<int>("nbrbase", ng + 1);
    ORAC_HIP(hipMemsetAsync(L.rcount, 0, sizeof(int) * ng, st));
    ORAC_HIP(hipMemsetAsync(L.rbase, 0, sizeof(int) * (ng + 1), st));
    // neighbour-side force slots: (list entries / 2) x atoms per group, with room to grow (size_list)
    const double dens = ng / S.volume, rl = d.rcut + d.rtol + d.rnei;
    const double pairs = 0.5 * ng * (4.0 / 3.0 * M_PI * rl * rl * rl * dens);
    size_half_slots(S, (long long)(1.5 * pairs * S.max_group) + 4096);
    p.half = 1;
    p.nb_mirror = L.mirror;
    p.nb_rcount = L.rcount;
    p.nb_rbase = L.rbase;
    p.part2 = B.get<double>("ppart2", (size_t)p.nblk * 9);
  }

  // ---- reciprocal space
  build_pme_plan(B, "", S.plan, d.nfft1, d.nfft2, d.nfft3, d.order, st);
  PmeDesc &m = S.pme;
  m = pme_desc_base(B, "", S.plan, n, S.oc, S.volume, d.alpha, 1.0e9, d.precision == 0, ctx->tuning, st);
  m.xp = S.xp;
  m.yp = S.yp;
  m.zp = S.zp;
  m.co = B.put("co", d.co, 9, st, false);
  m.q = q;
  m.fixscale = fix_scale_for(d.charge, n);
  m.fixinv = 1.0 / m.fixscale;
  // own force arrays: the reciprocal sum runs beside the direct sum and is added after both
  m.fx = B.get<double>("fxp", n);
  m.fy = B.get<double>("fyp", n);
  m.fz = B.get<double>("fzp", n);
  m.accumulate = 0;
  m.fx_dst = S.fx;
  m.fy_dst = S.fy;
  m.fz_dst = S.fz;
  m.atomp = d_ap;
  m.xpcm = xpcm;
  m.ypcm = ypcm;
  m.zpcm = zpcm;
  m.out = S.out;

  // ---- per-step reciprocal corrections (ferrf, iz = 1)
  CorrDesc &h = S.corr;
  std::memset(&h, 0, sizeof h);
  {
    const int nc = d.ncorr;
    S.corr_ia.resize(nc);
    S.corr_ib.resize(nc);
    S.corr_fud.assign(nc, 1.0);
    for (int k = 0; k < nc; ++k) {
      S.corr_ia[k] = d.corr_pairs[2 * k] - 1;
      S.corr_ib[k] = d.corr_pairs[2 * k + 1] - 1;
      ORAC_REQUIRE(S.corr_ia[k] >= 0 && S.corr_ia[k] < n && S.corr_ib[k] >= 0 && S.corr_ib[k] < n, E_ARG,
                   "correction pair out of range");
      if (d.corr_fudge) S.corr_fud[k] = d.corr_fudge[k];
    }
    corr_upload(S, nullptr, 0, st);
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
  // own force arrays, added with the reciprocal ones after the side stream joins
  h.fx = B.get<double>("fxc", n);
  h.fy = B.get<double>("fyc", n);
  h.fz = B.get<double>("fzc", n);
  h.write = 1;
  if (h.npairs > 0) {
    m.cfx = h.fx;
    m.cfy = h.fy;
    m.cfz = h.fz;
  }
  h.out = S.out + 36;

  // self energy (cself.f, kcut = 1e9): constant for fixed charges
  {
    double s[2] = {0.0, 0.0};
    for (int a = 0; a < n; ++a) s[d.ss_index[a] - 1] += d.charge[a] * d.charge[a];
    const double pi = 4.0 * std::atan(1.0);
    for (int k = 0; k < 2; ++k) S.self_e[k] = -s[k] * d.alpha / std::sqrt(pi);
  }
  ORAC_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------- batches

static void batch_upload(orac_hip_ctx *ctx, Batch &b) {
  hipStream_t st = ctx->stream;
  const size_t n = b.sys.size();
  b.hp.resize(n);
  b.hq.resize(n);
  b.hl.resize(n);
  b.hc.resize(n);
  b.any_corr = false;
  b.gens.resize(n);
  for (size_t k = 0; k < n; ++k) {
    System &S = *b.sys[k];
    b.gens[k] = S.gen;
    b.hp[k] = S.pme;
    pme_apply_tuning(b.hp[k], b.tu);
    b.hq[k] = S.pair;
    b.hl[k] = S.list;
    b.hl[k].cellorder = b.tu.cell_order ? 1 : 0;
    if (!b.tu.cell_order) {
      b.hq[k].perm = nullptr;  // owners and neighbour records in group order
      b.hq[k].soff = nullptr;
    }
    b.hc[k] = S.corr;
    if (S.corr.npairs > 0) b.any_corr = true;
  }
  // the energy-only lambda0 grids of alchemical systems follow the systems' own
  for (size_t k = 0; k < n; ++k)
    if (b.sys[k]->alch) {
      b.hp.push_back(b.sys[k]->pme0);
      pme_apply_tuning(b.hp.back(), b.tu);
    }
  b.shells = true;
  for (System *S : b.sys) b.shells = b.shells && S->rcutm > 0.0;
  b.hqm.clear();
  if (b.shells) {
    b.hqm.resize(n);
    for (size_t k = 0; k < n; ++k) {
      b.hqm[k] = b.sys[k]->pair_m;
      b.hqm[k].perm = b.hq[k].perm;
      b.hqm[k].soff = b.hq[k].soff;
    }
    b.dqm = b.B.put_vec("dqm", b.hqm, st, false);
  }
  b.dp = b.B.put_vec("dp", b.hp, st, false);
  b.dq = b.B.put_vec("dq", b.hq, st, false);
  b.dl = b.B.put_vec("dl", b.hl, st, false);
  b.dc = b.B.put_vec("dc", b.hc, st, false);
  ORAC_HIP(hipStreamSynchronize(st));
  for (int k = 0; k < 2; ++k) {
    if (b.exec[k]) ORAC_HIP(hipGraphExecDestroy(b.exec[k]));
    if (b.graph[k]) ORAC_HIP(hipGraphDestroy(b.graph[k]));
    b.exec[k] = nullptr;
    b.graph[k] = nullptr;
  }
  b.drop_md_graphs();
  b.md_gen = -1;
}

// descriptor copies of a batch follow its systems' buffers
static void ensure_fresh(orac_hip_ctx *ctx, Batch &b) {
  for (size_t k = 0; k < b.sys.size(); ++k)
    if (b.gens[k] != b.sys[k]->gen) {
      batch_upload(ctx, b);
      return;
    }
}

static int batch_create(orac_hip_ctx *ctx, const int32_t *ids, int n) {
  Engine &E = engine(ctx);
  ORAC_REQUIRE(n > 0 && ids, E_ARG, "a batch needs at least one system");
  auto b = std::make_unique<Batch>();
  b->tu = ctx->tuning;
  for (int k = 0; k < n; ++k) {
    System &S = get_sys(ctx, ids[k]);
    for (int j = 0; j < k; ++j) ORAC_REQUIRE(ids[j] != ids[k], E_ARG, "a system appears twice in the batch");
    if (k == 0) b->precision = S.precision;
    ORAC_REQUIRE(S.precision == b->precision, E_ARG, "all systems of a batch must have the same precision");
    b->ids.push_back(ids[k]);
    b->sys.push_back(&S);
  }
  for (auto &e : b->ev) ORAC_HIP(hipEventCreate(&e));
  ORAC_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  ORAC_HIP(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
  batch_upload(ctx, *b);
  const int id = E.next_batch++;
  E.batches[id] = std::move(b);
  return id;
}

// the stages of one force evaluation of every system in the batch
static void batch_enqueue(orac_hip_ctx *ctx, Batch &b, bool rebuild, bool events, bool md_lists = false) {
  hipStream_t st = ctx->stream;
  const int n = (int)b.sys.size();
  const int npme = (int)b.hp.size();
  auto mark = [&](int k) {
    if (events) ORAC_HIP(hipEventRecord(b.ev[k], st));
  };
  auto pme = [&](hipStream_t s) {
    const hipEvent_t *pev = events ? &b.ev[ST_P0] : nullptr;
    if (b.precision == 0)
      launch_pme<double>(b.dp, b.hp.data(), npme, s, b.tu, pev);
    else
      launch_pme<float>(b.dp, b.hp.data(), npme, s, b.tu, pev);
  };
  mark(ST_BEGIN);
  launch_centres(b.dq, b.hq.data(), n, st);
  mark(ST_CENTRES);
  // The reciprocal sum needs the coordinates and the molecule centres only:
  // unless the stages are being timed it runs on the side stream beside the
  // list build and the pair kernel (both latency bound), and joins before
  // its forces are added.
  const bool side = !events && ctx->side;
  if (side) {
    ORAC_HIP(hipEventRecord(b.fork, st));
    ORAC_HIP(hipStreamWaitEvent(ctx->side, b.fork, 0));
    pme(ctx->side);
    if (b.any_corr) launch_corr(b.dc, b.hc.data(), n, ctx->side);
    ORAC_HIP(hipEventRecord(b.join, ctx->side));
  }
  if (md_lists) {  // every system on its own update schedule (device step counters)
    launch_list_build(b.dl_md, b.hl_md.data(), n, LIST_COUNT | LIST_WRITE, st);
    if (b.shells) launch_list_inner(b.dl_md, b.hl_md.data(), n, st);
  } else if (rebuild) {
    launch_list_build(b.dl, b.hl.data(), n, LIST_COUNT | LIST_WRITE, st);
    if (b.shells) launch_list_inner(b.dl, b.hl.data(), n, st);
  }
  mark(ST_LIST);
  hipEvent_t kb = events ? b.ev[ST_KBEG] : nullptr, ke = events ? b.ev[ST_KEND] : nullptr;
  if (b.precision == 0)
    launch_pairs<double>(b.dq, b.hq.data(), n, st, b.tu, kb, ke);
  else
    launch_pairs<float>(b.dq, b.hq.data(), n, st, b.tu, kb, ke);
  mark(ST_PAIRS);
  if (side) {
    ORAC_HIP(hipStreamWaitEvent(st, b.join, 0));
    mark(ST_PME);
  } else {
    pme(st);
    mark(ST_PME);
    if (b.any_corr) launch_corr(b.dc, b.hc.data(), n, st);
  }
  launch_pme_combine(b.dp, b.hp.data(), npme, st);
  mark(ST_END);
}

// The inner shell of every system (systems with two shells): centres, then the
// pair kernel on the inner-shell descriptors, on the list of the last build.
static void batch_enqueue_mid(orac_hip_ctx *ctx, Batch &b) {
  hipStream_t st = ctx->stream;
  const int n = (int)b.sys.size();
  launch_centres(b.dq, b.hq.data(), n, st);
  if (b.precision == 0)
    launch_pairs<double>(b.dqm, b.hqm.data(), n, st, b.tu, nullptr, nullptr);
  else
    launch_pairs<float>(b.dqm, b.hqm.data(), n, st, b.tu, nullptr, nullptr);
}

static void batch_eval(orac_hip_ctx *ctx, Batch &b, bool rebuild, bool use_graph) {
  ensure_fresh(ctx, b);
  for (System *S : b.sys)
    if (!S->have_list || S->list_fmt != b.tu.cell_order) rebuild = true;
  hipStream_t st = ctx->stream;
  if (use_graph && !b.profile) {
    const int k = rebuild ? 1 : 0;
    if (!b.exec[k]) {
      ORAC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        batch_enqueue(ctx, b, rebuild, false);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      ORAC_HIP(hipStreamEndCapture(st, &b.graph[k]));
      ORAC_HIP(hipGraphInstantiate(&b.exec[k], b.graph[k], nullptr, nullptr, 0));
    }
    ORAC_HIP(hipGraphLaunch(b.exec[k], st));
  } else {
    batch_enqueue(ctx, b, rebuild, b.profile);
  }
  ORAC_HIP(hipGetLastError());
  if (rebuild)
    for (System *S : b.sys) {
      S->have_list = true;
      S->list_fmt = b.tu.cell_order;
    }
}

static int solo_batch(orac_hip_ctx *ctx, int32_t sys) {
  System &S = get_sys(ctx, sys);
  Engine &E = engine(ctx);
  if (!S.solo || !E.batches.count(S.solo)) S.solo = batch_create(ctx, &sys, 1);
  return S.solo;
}

// Size the neighbour list from the counts at the current coordinates: room
// per owner for 1.5 x the largest count + 64 (at least the creation estimate),
// so that density fluctuations during dynamics stay within the slab.  Clears
// the sticky overflow flag.  Run whenever the host sets coordinates.
static void size_list(orac_hip_ctx *ctx, int32_t id) {
  System &S = get_sys(ctx, id);
  Batch &b = get_batch(ctx, solo_batch(ctx, id));
  ensure_fresh(ctx, b);
  hipStream_t st = ctx->stream;
  launch_centres(b.dq, b.hq.data(), 1, st);
  launch_list_build(b.dl, b.hl.data(), 1, LIST_COUNT, st);
  int tot[4] = {0, 0, 0, 0};
  ORAC_HIP(hipMemcpyAsync(tot, S.list.total, sizeof tot, hipMemcpyDeviceToHost, st));
  ORAC_HIP(hipStreamSynchronize(st));
  const int ng = S.ngrp;
  const int need = std::min(ng, ((int)(1.5 * tot[2]) + 64 + 31) / 32 * 32);
  S.have_list = false;  // the count pass rewrote the counts
  launch_zero_u64(reinterpret_cast<unsigned long long *>(S.list.total + 2), 1, st);
  if (need > S.slab) {
    ORAC_REQUIRE((size_t)ng * (size_t)need < (size_t)INT32_MAX, E_CAPACITY,
                 "neighbour list larger than 2^31 entries");
    S.slab = need;
    S.list_cap = (size_t)ng * (size_t)need;
    S.list.out = S.B.get<int>("nbidx", S.list_cap);
    S.list.cap = (long long)S.list_cap;
    S.list.slab = need;
    S.pair.nb_idx = S.list.out;
    S.pair.nb_slab = need;
    if (S.list.out_m) {  // the inner list follows the slab
      S.list.out_m = S.B.get<int>("nbidx_m", S.list_cap);
      S.pair_m.nb_idx = S.list.out_m;
      S.pair_m.nb_slab = need;
    }
    if (S.list.split) {
      S.list.rout = S.B.get<int>("nbrout", S.list_cap);
      S.list.mirror = S.B.get<int>("nbmirror", S.list_cap);
      S.pair.nb_mirror = S.list.mirror;
    }
    S.gen++;
  }
  // the count pass counted both directions: slots for half of them, each group pair's upper-group atoms
  if (S.list.split) size_half_slots(S, (long long)(0.75 * tot[0] * S.max_group) + 4096);
}

// Systems of b whose neighbour list overflowed its slab at some build since
// their list was last sized (the sticky flag total[3]); synchronises.
static std::vector<int> list_overflows(orac_hip_ctx *ctx, Batch &b) {
  const size_t n = b.sys.size();
  std::vector<int> flag(n, 0), bad;
  for (size_t k = 0; k < n; ++k)
    ORAC_HIP(hipMemcpyAsync(&flag[k], b.sys[k]->list.total + 3, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  ORAC_HIP(hipStreamSynchronize(ctx->stream));
  for (size_t k = 0; k < n; ++k)
    if (flag[k]) bad.push_back(b.ids[k]);
  return bad;
}

// time of one evaluation (ms, mean over reps after one warm-up) of a batch
static double batch_time(orac_hip_ctx *ctx, Batch &b, int reps, bool rebuild) {
  hipStream_t st = ctx->stream;
  hipEvent_t e0, e1;
  ORAC_HIP(hipEventCreate(&e0));
  ORAC_HIP(hipEventCreate(&e1));
  float ms = 0.f;
  try {
    batch_eval(ctx, b, rebuild, true);
    ORAC_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) batch_eval(ctx, b, rebuild, true);
    ORAC_HIP(hipEventRecord(e1, st));
    ORAC_HIP(hipEventSynchronize(e1));
    ORAC_HIP(hipEventElapsedTime(&ms, e0, e1));
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  ORAC_HIP(hipEventDestroy(e0));
  ORAC_HIP(hipEventDestroy(e1));
  return ms / std::max(1, reps);
}

// ---------------------------------------------------------------- molecular dynamics

static void md_setup(orac_hip_ctx *ctx, System &S, int nrigid, const int32_t *rigid_first, const double *rigid_len,
                     int nterm, const int32_t *terms, const double *tk, const double *tr0) {
  hipStream_t st = ctx->stream;
  BufSet &B = S.B;
  const int n = S.natoms;
  ORAC_REQUIRE(nrigid >= 0 && nterm >= 0, E_ARG, "negative counts");
  ORAC_REQUIRE((nrigid == 0 || (rigid_first && rigid_len)) && (nterm == 0 || (terms && tk && tr0)), E_ARG,
               "null array argument");
  // units: rigid triplets of consecutive atoms, every other atom alone
  std::vector<int> rig_at(n, -1);
  for (int r = 0; r < nrigid; ++r) {
    const int a = rigid_first[r] - 1;
    ORAC_REQUIRE(a >= 0 && a + 2 < n, E_ARG, "rigid unit out of range");
    for (int k = 0; k < 3; ++k) {
      ORAC_REQUIRE(rig_at[a + k] < 0, E_ARG, "an atom is in two rigid units");
      rig_at[a + k] = r;
    }
    for (int k = 0; k < 3; ++k) ORAC_REQUIRE(rigid_len[3 * r + k] > 0.0, E_ARG, "rigid lengths must be positive");
  }
  std::vector<int> ufirst;
  std::vector<unsigned char> urigid;
  std::vector<double> ud2;
  for (int a = 0; a < n;) {
    ufirst.push_back(a);
    if (rig_at[a] >= 0) {
      ORAC_REQUIRE(rigid_first[rig_at[a]] - 1 == a, E_STATE, "rigid unit bookkeeping");
      urigid.push_back(1);
      for (int k = 0; k < 3; ++k) ud2.push_back(rigid_len[3 * rig_at[a] + k] * rigid_len[3 * rig_at[a] + k]);
      a += 3;
    } else {
      urigid.push_back(0);
      for (int k = 0; k < 3; ++k) ud2.push_back(0.0);
      a += 1;
    }
  }
  // harmonic terms and their per-atom gather
  std::vector<int> bi(std::max(1, nterm)), bj(std::max(1, nterm));
  std::vector<double> bk(std::max(1, nterm)), br0(std::max(1, nterm));
  std::vector<std::vector<int>> per(n);
  for (int t = 0; t < nterm; ++t) {
    bi[t] = terms[2 * t] - 1;
    bj[t] = terms[2 * t + 1] - 1;
    ORAC_REQUIRE(bi[t] >= 0 && bi[t] < n && bj[t] >= 0 && bj[t] < n && bi[t] != bj[t], E_ARG,
                 "harmonic term out of range");
    bk[t] = tk[t];
    br0[t] = tr0[t];
    per[bi[t]].push_back(t + 1);
    per[bj[t]].push_back(-(t + 1));
  }
  std::vector<int> goff(n + 1, 0), gt;
  for (int a = 0; a < n; ++a) {
    for (int v : per[a]) gt.push_back(v);
    goff[a + 1] = (int)gt.size();
  }
  if (gt.empty()) gt.push_back(0);
  {
    std::vector<double> m(n);
    ORAC_HIP(hipMemcpy(m.data(), S.pair.mass, n * sizeof(double), hipMemcpyDeviceToHost));
    for (int a = 0; a < n; ++a) ORAC_REQUIRE(m[a] > 0.0, E_ARG, "dynamics needs positive masses");
  }
  MdDesc &d = S.md;
  std::memset(&d, 0, sizeof d);
  d.natoms = n;
  d.nunits = (int)ufirst.size();
  d.nbond = nterm;
  d.x = B.get<double>("md_x", n);
  d.y = B.get<double>("md_y", n);
  d.z = B.get<double>("md_z", n);
  d.vx = B.get<double>("md_vx", n);
  d.vy = B.get<double>("md_vy", n);
  d.vz = B.get<double>("md_vz", n);
  d.xp = S.xp;
  d.yp = S.yp;
  d.zp = S.zp;
  d.fx = S.fx;
  d.fy = S.fy;
  d.fz = S.fz;
  d.bfx = B.get<double>("md_bfx", n);
  d.bfy = B.get<double>("md_bfy", n);
  d.bfz = B.get<double>("md_bfz", n);
  d.mass = S.pair.mass;
  d.fmx = S.rcutm > 0.0 ? S.fmx : nullptr;
  d.fmy = S.rcutm > 0.0 ? S.fmy : nullptr;
  d.fmz = S.rcutm > 0.0 ? S.fmz : nullptr;
  d.ufirst = B.put_vec("md_ufirst", ufirst, st, false);
  d.urigid = B.put_vec("md_urigid", urigid, st, false);
  d.ud2 = B.put_vec("md_ud2", ud2, st, false);
  d.bi = B.put_vec("md_bi", bi, st, false);
  d.bj = B.put_vec("md_bj", bj, st, false);
  d.bk = B.put_vec("md_bk", bk, st, false);
  d.br0 = B.put_vec("md_br0", br0, st, false);
  d.bpf = B.get<double>("md_bpf", 3 * (size_t)std::max(1, nterm));
  d.bpe = B.get<double>("md_bpe", (size_t)std::max(1, nterm));
  d.bg_off = B.put_vec("md_bgoff", goff, st, false);
  d.bg_term = B.put_vec("md_bgterm", gt, st, false);
  for (int k = 0; k < 9; ++k) d.oc[k] = S.oc[k];
  d.out = B.get<double>("md_out", 8);
  d.fail = B.get<int>("md_fail", 2);
  d.step = B.get<int>("md_step", 2);
  launch_zero_u64(reinterpret_cast<unsigned long long *>(d.step), 1, st);
  for (double *p : {d.vx, d.vy, d.vz, d.bfx, d.bfy, d.bfz, d.out}) launch_zero_f64(p, p == d.out ? 8 : n, st);
  launch_zero_u64(reinterpret_cast<unsigned long long *>(d.fail), 1, st);
  // positions from the current reduced coordinates: x = co xp
  {
    std::vector<double> xp(n), yp(n), zp(n), x(n), y(n), z(n);
    ORAC_HIP(hipMemcpyAsync(xp.data(), S.xp, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(yp.data(), S.yp, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(zp.data(), S.zp, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    const double *c = S.co;
    for (int a = 0; a < n; ++a) {
      x[a] = c[0] * xp[a] + c[3] * yp[a] + c[6] * zp[a];
      y[a] = c[1] * xp[a] + c[4] * yp[a] + c[7] * zp[a];
      z[a] = c[2] * xp[a] + c[5] * yp[a] + c[8] * zp[a];
    }
    ORAC_HIP(hipMemcpyAsync(d.x, x.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    ORAC_HIP(hipMemcpyAsync(d.y, y.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    ORAC_HIP(hipMemcpyAsync(d.z, z.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  }
  S.md_d = B.put("md_desc", &d, 1, st, false);
  ORAC_HIP(hipStreamSynchronize(st));
  S.md_on = true;
  S.md_primed = false;
  engine(ctx).md_gen++;
}

static void md_upload(orac_hip_ctx *ctx, Batch &b, int list_every) {
  ensure_fresh(ctx, b);
  const long gen = engine(ctx).md_gen;
  if (b.md_gen == gen && b.dm && b.md_every == list_every) return;
  const size_t n = b.sys.size();
  b.hm.resize(n);
  b.hl_md.resize(n);
  for (size_t k = 0; k < n; ++k) {
    ORAC_REQUIRE(b.sys[k]->md_on, E_STATE, "dynamics is not set up for every system of the batch");
    b.hm[k] = b.sys[k]->md;
    b.hl_md[k] = b.hl[k];
    b.hl_md[k].step = b.sys[k]->md.step;
    b.hl_md[k].every = list_every;
  }
  b.dm = b.B.put_vec("dm", b.hm, ctx->stream, false);
  b.dl_md = b.B.put_vec("dl_md", b.hl_md, ctx->stream, false);
  ORAC_HIP(hipStreamSynchronize(ctx->stream));
  b.drop_md_graphs();
  b.md_gen = gen;
  b.md_every = list_every;
}

// Forces at the current positions (fast terms and the full nonbonded
// evaluation with a fresh list) for the systems of b that lack them.  They
// are computed in a batch of their own, so systems already integrating keep
// their forces and a system's trajectory does not depend on its batch-mates.
static void md_prime(orac_hip_ctx *ctx, Batch &b) {
  std::vector<int32_t> ids;
  for (size_t k = 0; k < b.sys.size(); ++k)
    if (!b.sys[k]->md_primed) ids.push_back(b.ids[k]);
  if (ids.empty()) return;
  Engine &E = engine(ctx);
  const int pid = batch_create(ctx, ids.data(), (int)ids.size());
  Batch &p = *E.batches[pid];
  try {
    p.tu = b.tu;
    batch_upload(ctx, p);
    md_upload(ctx, p, 1);
    launch_md_fast_forces(p.dm, p.hm.data(), (int)ids.size(), ctx->stream);
    batch_eval(ctx, p, true, false);
    if (p.shells) batch_enqueue_mid(ctx, p);
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
  } catch (...) {
    E.batches.erase(pid);
    throw;
  }
  for (System *S : p.sys) S->md_primed = true;
  E.batches.erase(pid);
}

// one outer r-RESPA step of every system in the batch (see md.hip)
static void md_enqueue_step(orac_hip_ctx *ctx, Batch &b, double dt, int nmid, int ninner) {
  hipStream_t st = ctx->stream;
  const int n = (int)b.sys.size();
  if (nmid <= 0) {
    const double h = dt / ninner;
    // slow half kick, then the first fast substep's drift
    launch_md_units(b.dm, b.hm.data(), n, dt, 0.0, h, MD_KICK_SLOW | MD_VERLET, st);
    launch_md_fast_forces(b.dm, b.hm.data(), n, st);
    // end of one fast substep (half kick), start of the next (drift)
    for (int k = 1; k < ninner; ++k) {
      launch_md_units(b.dm, b.hm.data(), n, dt, 0.0, h, MD_KICK_FAST | MD_VERLET, st);
      launch_md_fast_forces(b.dm, b.hm.data(), n, st);
    }
    batch_enqueue(ctx, b, true, false, true);
    // end of the last fast substep, then the slow half kick
    launch_md_units(b.dm, b.hm.data(), n, dt, 0.0, h, MD_KICK_FAST | MD_KICK_SLOW, st);
  } else {
    // two nonbonded shells, three levels in mtsmd.f's order: the outer shell
    // with PME and corrections over dt, nmid steps of the inner shell over
    // hm = dt/nmid, each with ninner fast substeps over h = hm/ninner
    const double hm = dt / nmid, h = hm / ninner;
    launch_md_units(b.dm, b.hm.data(), n, dt, hm, h, MD_KICK_SLOW, st);
    for (int im = 0; im < nmid; ++im) {
      launch_md_units(b.dm, b.hm.data(), n, dt, hm, h, MD_KICK_MID | MD_VERLET, st);
      launch_md_fast_forces(b.dm, b.hm.data(), n, st);
      for (int k = 1; k < ninner; ++k) {
        launch_md_units(b.dm, b.hm.data(), n, dt, hm, h, MD_KICK_FAST | MD_VERLET, st);
        launch_md_fast_forces(b.dm, b.hm.data(), n, st);
      }
      launch_md_units(b.dm, b.hm.data(), n, dt, hm, h, MD_KICK_FAST, st);
      if (im == nmid - 1) batch_enqueue(ctx, b, true, false, true);  // the outer shell at the step's end
      batch_enqueue_mid(ctx, b);
      launch_md_units(b.dm, b.hm.data(), n, dt, hm, h, MD_KICK_MID, st);
    }
    launch_md_units(b.dm, b.hm.data(), n, dt, hm, h, MD_KICK_SLOW, st);
  }
  launch_md_tick(b.dm, n, st);
}

static void md_run(orac_hip_ctx *ctx, Batch &b, int nsteps, double dt, int nmid, int ninner, int list_every,
                   bool use_graph) {
  ORAC_REQUIRE(nsteps >= 0 && dt > 0.0 && ninner >= 1 && list_every >= 1 && nmid >= 0, E_ARG,
               "bad integration parameters");
  ORAC_REQUIRE(nmid == 0 || b.shells, E_STATE, "an inner-shell level needs two shells on every system");
  ORAC_REQUIRE(nmid > 0 || !b.shells, E_STATE, "systems with two shells integrate with an inner-shell level");
  md_upload(ctx, b, list_every);
  md_prime(ctx, b);
  // lists built in another layout are rebuilt (and the forces with them)
  bool fmt = false;
  for (System *S : b.sys) fmt = fmt || S->list_fmt != b.tu.cell_order;
  if (fmt) {
    launch_md_fast_forces(b.dm, b.hm.data(), (int)b.sys.size(), ctx->stream);
    batch_eval(ctx, b, true, false);
    if (b.shells) batch_enqueue_mid(ctx, b);
  }
  if (b.md_dt != dt || b.md_ninner != ninner || b.md_nmid != nmid) {
    b.drop_md_graphs();
    b.md_dt = dt;
    b.md_ninner = ninner;
    b.md_nmid = nmid;
  }
  hipStream_t st = ctx->stream;
  for (int s = 0; s < nsteps; ++s) {
    if (use_graph) {
      if (!b.mexec) {
        ORAC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
          md_enqueue_step(ctx, b, dt, nmid, ninner);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        ORAC_HIP(hipStreamEndCapture(st, &b.mgraph));
        ORAC_HIP(hipGraphInstantiate(&b.mexec, b.mgraph, nullptr, nullptr, 0));
      }
      ORAC_HIP(hipGraphLaunch(b.mexec, st));
    } else {
      md_enqueue_step(ctx, b, dt, nmid, ninner);
    }
  }
  for (System *S : b.sys) S->have_list = true;
  ORAC_HIP(hipGetLastError());
  // a list cut short by its slab leaves that step's forces wrong: fail loudly
  const std::vector<int> bad = list_overflows(ctx, b);
  ORAC_REQUIRE(bad.empty(), E_CAPACITY,
               "neighbour list of system " + std::to_string(bad.empty() ? -1 : bad[0]) +
                   " outgrew its slab during dynamics; its trajectory is invalid from that step on "
                   "(setting its positions again resizes the list)");
}

}  // namespace orac

using namespace orac;

// ================================================================ C ABI: systems

extern "C" int orac_hip_system_create(orac_hip_ctx *ctx, const orac_hip_system_desc *desc, int32_t *sys) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(desc && sys, E_ARG, "null argument");
    auto S = std::make_unique<System>();
    system_build(ctx, *S, *desc);
    Engine &E = engine(ctx);
    const int id = E.next_sys++;
    E.systems[id] = std::move(S);
    *sys = id;
  });
}

extern "C" int orac_hip_system_destroy(orac_hip_ctx *ctx, int32_t sys) {
  return guarded([&] {
    set_device(ctx);
    get_sys(ctx, sys);
    Engine &E = engine(ctx);
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
    // batches holding the system go with it
    for (auto it = E.batches.begin(); it != E.batches.end();) {
      const auto &ids = it->second->ids;
      if (std::find(ids.begin(), ids.end(), sys) != ids.end())
        it = E.batches.erase(it);
      else
        ++it;
    }
    E.systems.erase(sys);
  });
}

extern "C" int orac_hip_system_bytes(orac_hip_ctx *ctx, int32_t sys, int64_t *bytes) {
  return guarded([&] {
    ORAC_REQUIRE(bytes, E_ARG, "null output");
    *bytes = (int64_t)get_sys(ctx, sys).B.bytes();
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
    size_list(ctx, sys);
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
    // replayed from the system's own HIP graph (eager while profiling)
    Batch &b = get_batch(ctx, solo_batch(ctx, sys));
    batch_eval(ctx, b, rebuild_list != 0, true);
    if (b.shells) batch_enqueue_mid(ctx, b);
  });
}

extern "C" int orac_hip_system_profile(orac_hip_ctx *ctx, int32_t sys, int32_t enable) {
  return guarded([&] {
    set_device(ctx);
    get_batch(ctx, solo_batch(ctx, sys)).profile = enable != 0;
  });
}

static void stage_ms(Batch &b, float *ms, int32_t n) {
  ORAC_REQUIRE(ms && n >= 0, E_ARG, "bad buffer");
  ORAC_REQUIRE(b.profile, E_STATE, "profiling is not enabled");
  ORAC_HIP(hipEventSynchronize(b.ev[ST_END]));
  for (int k = 0; k < n && k < NSTAGE_MS; ++k) ORAC_HIP(hipEventElapsedTime(&ms[k], b.ev[k], b.ev[k + 1]));
  if (n > NSTAGE_MS) ORAC_HIP(hipEventElapsedTime(&ms[NSTAGE_MS], b.ev[ST_KBEG], b.ev[ST_KEND]));
  // the reciprocal sum's spread chain, FFT stage (both transforms and the convolution) and gather
  for (int k = 0; k < 3 && NSTAGE_MS + 1 + k < n; ++k)
    ORAC_HIP(hipEventElapsedTime(&ms[NSTAGE_MS + 1 + k], b.ev[ST_P0 + k], b.ev[ST_P1 + k]));
}

extern "C" int orac_hip_system_stage_ms(orac_hip_ctx *ctx, int32_t sys, float *ms, int32_t n) {
  return guarded([&] {
    set_device(ctx);
    stage_ms(get_batch(ctx, solo_batch(ctx, sys)), ms, n);
  });
}

extern "C" int orac_hip_system_results(orac_hip_ctx *ctx, int32_t sys, double *fx, double *fy, double *fz,
                                       double *energies) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    hipStream_t st = ctx->stream;
    double out[48], outm[16];
    int tot[2];
    ORAC_HIP(hipMemcpyAsync(out, S.out, sizeof out, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(tot, S.list.total, sizeof tot, hipMemcpyDeviceToHost, st));
    size_t nb = (size_t)S.natoms * sizeof(double);
    if (fx) ORAC_HIP(hipMemcpyAsync(fx, S.fx, nb, hipMemcpyDeviceToHost, st));
    if (fy) ORAC_HIP(hipMemcpyAsync(fy, S.fy, nb, hipMemcpyDeviceToHost, st));
    if (fz) ORAC_HIP(hipMemcpyAsync(fz, S.fz, nb, hipMemcpyDeviceToHost, st));
    std::vector<double> mx, my, mz;
    const bool two = S.rcutm > 0.0;
    if (two) {  // the inner shell's forces and energies join the outer ones
      ORAC_HIP(hipMemcpyAsync(outm, S.out_m, sizeof outm, hipMemcpyDeviceToHost, st));
      mx.resize(S.natoms);
      my.resize(S.natoms);
      mz.resize(S.natoms);
      ORAC_HIP(hipMemcpyAsync(mx.data(), S.fmx, nb, hipMemcpyDeviceToHost, st));
      ORAC_HIP(hipMemcpyAsync(my.data(), S.fmy, nb, hipMemcpyDeviceToHost, st));
      ORAC_HIP(hipMemcpyAsync(mz.data(), S.fmz, nb, hipMemcpyDeviceToHost, st));
    }
    ORAC_HIP(hipStreamSynchronize(st));
    if (two) {
      for (int a = 0; a < S.natoms; ++a) {
        if (fx) fx[a] += mx[a];
        if (fy) fy[a] += my[a];
        if (fz) fz[a] += mz[a];
      }
      for (int k = 0; k < 15; ++k) out[20 + k] += outm[k];
    }
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
      // corrections; the intra-alchemical term joins them (mtsmd.f:3041: coul_bnd_slt += falch_slt)
      energies[7] = S.nqq > 0 ? out[36 + 0] + out[36 + 4] : out[36 + 0];
      energies[8] = S.nqq > 0 ? out[36 + 1] + out[36 + 5] : out[36 + 1];
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

// ================================================================ C ABI: alchemy

// charges, soft-core factors and self terms at the system's lambdas
static void alch_apply_lambda(orac_hip_ctx *ctx, System &S) {
  hipStream_t st = ctx->stream;
  const int n = S.natoms;
  double fac[72];
  alch_factors(S.lam, S.intramol, S.ljintra, S.qsoft, S.ljsoft, S.shifted, fac);
  std::vector<double> qc(n), q0(n);
  alch_charges(n, S.q_host.data(), S.ss_alch.data(), S.lam[4], S.lam[5], qc.data());
  alch_charges(n, S.q_host.data(), S.ss_alch.data(), S.lam[6], S.lam[7], q0.data());
  ORAC_HIP(hipMemcpyAsync(S.afac, fac, sizeof fac, hipMemcpyHostToDevice, st));
  ORAC_HIP(hipMemcpyAsync(S.qc, qc.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  ORAC_HIP(hipMemcpyAsync(S.q0, q0.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  if (S.nqq > 0) {
    // per-atom lambdaq: +laddq added, -lrmq removed, 0 otherwise (mtsmd.f:746-757)
    std::vector<double> lq(n), lq0(n);
    for (int a = 0; a < n; ++a) {
      const int k = S.ss_alch[a];
      lq[a] = k == 2 ? S.lam[4] : (k == -2 ? -S.lam[5] : 0.0);
      lq0[a] = k == 2 ? S.lam[6] : (k == -2 ? -S.lam[7] : 0.0);
    }
    ORAC_HIP(hipMemcpyAsync(S.lq, lq.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    ORAC_HIP(hipMemcpyAsync(S.lq0, lq0.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    ORAC_HIP(hipStreamSynchronize(st));
  }
  // cself_alchemy.f: solute atoms only, kcut = 1e9 (erfc term 0)
  double fs = 0.0, en0 = 0.0;
  for (int a = 0; a < n; ++a)
    if (S.ss_host[a] == 1) {
      fs = fs + qc[a] * qc[a];
      en0 = en0 + q0[a] * q0[a];
    }
  const double pi = 4.0 * std::atan(1.0);
  const double e = erfc_as(1.0e9 / (2.0 * S.alpha));
  fs = -fs * S.alpha / std::sqrt(pi) * (1.0 - e);
  en0 = -en0 * S.alpha / std::sqrt(pi) * (1.0 - e);
  S.self_e[0] = fs;
  S.self_e[1] = 0.0;
  S.self_dw = fs - en0;
  ORAC_HIP(hipStreamSynchronize(st));
}

extern "C" int orac_hip_system_set_alchemy(orac_hip_ctx *ctx, int32_t sys, const orac_hip_alchemy_desc *a) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_REQUIRE(a && a->ss_alchemy && a->rmin && a->epsm, E_ARG, "null alchemy description");
    ORAC_REQUIRE(!S.alch, E_STATE, "the system is already alchemical");
    hipStream_t st = ctx->stream;
    BufSet &B = S.B;
    const int n = S.natoms;
    AlchTables A = alch_tables(n, a->ss_alchemy, a->rmin, a->epsm, a->opls);
    S.ss_alch.assign(a->ss_alchemy, a->ss_alchemy + n);
    S.intramol = a->intramol != 0;
    S.ljintra = a->ljintra != 0;
    S.shifted = a->shifted != 0;
    S.qsoft = S.shifted ? a->qshift : a->qsoft;
    S.ljsoft = S.shifted ? a->ljshift : a->ljsoft;
    S.pair.alch_shifted = S.shifted ? 1 : 0;
    // the soft-core pairs walk the full list (k_pairs): leave the split layout
    S.list.split = 0;
    S.pair.half = 0;
    S.have_list = false;
    S.pair.type = B.put_vec("atype", A.cls, st, false);
    S.pair.ntypes = A.ncls;
    S.pair.alch_pair = B.put_vec("apair", A.pair, st, false);
    S.afac = B.get<double>("afac", 72);
    S.pair.alch_fac = S.afac;
    S.qc = B.get<double>("qc", n);
    S.q0 = B.get<double>("q0", n);
    S.pme.q = S.qc;
    S.corr.q = S.qc;
    S.corr.q0 = S.q0;
    if (a->nqq > 0) {
      // ferrf_alchemy_add on every evaluation (mtsmd.f:3034-3040), in the
      // system's correction launch after its ferrf pairs
      ORAC_REQUIRE(a->listqq, E_ARG, "null intra-alchemical pair list");
      corr_upload(S, a->listqq, a->nqq, st);
      S.lq = B.get<double>("lq", n);
      S.lq0 = B.get<double>("lq0", n);
      S.corr.qa = B.put("qa", S.q_host.data(), n, st, false);
      S.corr.lq = S.lq;
      S.corr.lq0 = S.lq0;
      S.corr.intramol = S.intramol ? 1 : 0;
    }
    PmeDesc &e = S.pme0;
    e = pme_desc_base(B, "a0_", S.plan, n, S.oc, S.volume, S.alpha, 1.0e9, S.precision == 0, ctx->tuning, st);
    e.xp = S.pme.xp;
    e.yp = S.pme.yp;
    e.zp = S.pme.zp;
    e.co = S.pme.co;
    e.q = S.q0;
    e.fixscale = S.pme.fixscale;
    e.fixinv = S.pme.fixinv;
    e.fx = S.pme.fx;  // never written (energy only)
    e.fy = S.pme.fy;
    e.fz = S.pme.fz;
    e.accumulate = 0;
    e.fx_dst = e.fy_dst = e.fz_dst = nullptr;
    e.atomp = S.pme.atomp;
    e.xpcm = S.pme.xpcm;
    e.ypcm = S.pme.ypcm;
    e.zpcm = S.pme.zpcm;
    e.out = S.out + 48;
    e.energy_only = 1;
    S.alch = true;
    alch_apply_lambda(ctx, S);
    S.gen++;
  });
}

extern "C" int orac_hip_system_set_lambda(orac_hip_ctx *ctx, int32_t sys, const double *lam) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_REQUIRE(S.alch, E_STATE, "the system is not alchemical (orac_hip_system_set_alchemy)");
    ORAC_REQUIRE(lam, E_ARG, "null lambda array");
    for (int k = 0; k < 8; ++k) {
      ORAC_REQUIRE(std::isfinite(lam[k]) && std::fabs(lam[k]) <= 1.0, E_ARG, "lambdas must lie in [-1, 1]");
      S.lam[k] = lam[k];
    }
    alch_apply_lambda(ctx, S);
  });
}

extern "C" int orac_hip_system_alchemy_results(orac_hip_ctx *ctx, int32_t sys, double *out) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_REQUIRE(S.alch, E_STATE, "the system is not alchemical");
    ORAC_REQUIRE(out, E_ARG, "null output");
    double o[72];
    ORAC_HIP(hipMemcpyAsync(o, S.out, sizeof o, hipMemcpyDeviceToHost, ctx->stream));
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
    double corr = o[36] + o[37] - o[38];
    if (S.nqq > 0) corr += o[40] + o[41] - o[42];  // ferrf_alchemy_add.f:203: dwrk_al
    const double direct = o[35], recip = o[0] - o[48];
    out[1] = direct;
    out[2] = recip;
    out[3] = corr;
    out[4] = S.self_dw;
    out[0] = direct + recip + corr + S.self_dw;
    out[5] = o[48];
  });
}

// ================================================================ C ABI: batches

extern "C" int orac_hip_batch_create(orac_hip_ctx *ctx, int32_t nsys, const int32_t *systems, int32_t *batch) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(batch, E_ARG, "null output");
    *batch = batch_create(ctx, systems, nsys);
  });
}

extern "C" int orac_hip_batch_destroy(orac_hip_ctx *ctx, int32_t batch) {
  return guarded([&] {
    set_device(ctx);
    get_batch(ctx, batch);
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
    engine(ctx).batches.erase(batch);
  });
}

extern "C" int orac_hip_batch_eval(orac_hip_ctx *ctx, int32_t batch, int32_t rebuild_list, int32_t use_graph) {
  return guarded([&] {
    set_device(ctx);
    Batch &b = get_batch(ctx, batch);
    batch_eval(ctx, b, rebuild_list != 0, use_graph != 0);
    if (b.shells) batch_enqueue_mid(ctx, b);
  });
}

extern "C" int orac_hip_batch_profile(orac_hip_ctx *ctx, int32_t batch, int32_t enable) {
  return guarded([&] { get_batch(ctx, batch).profile = enable != 0; });
}

extern "C" int orac_hip_batch_stage_ms(orac_hip_ctx *ctx, int32_t batch, float *ms, int32_t n) {
  return guarded([&] {
    set_device(ctx);
    stage_ms(get_batch(ctx, batch), ms, n);
  });
}

static Tuning tuning_from(const int32_t *t, int32_t n, Tuning base) {
  if (n > 0) base.lds_elems = t[0];
  if (n > 1) base.pair_waves = t[1];
  if (n > 2) base.spread_tiled = t[2];
  if (n > 3) base.xcd_remap = t[3];
  if (n > 4) base.cell_order = t[4];
  if (n > 5) base.persistent = t[5];
  ORAC_REQUIRE(base.lds_elems >= 64 && base.lds_elems <= 2048 && (base.lds_elems & (base.lds_elems - 1)) == 0, E_ARG,
               "lds_elems must be a power of two in [64, 2048]");
  ORAC_REQUIRE(base.pair_waves >= 0 && base.pair_waves <= 4 && base.pair_waves != 1, E_ARG,
               "pair_waves must be 0 (default), 2, 3 or 4");
  ORAC_REQUIRE(base.spread_tiled >= 0 && base.spread_tiled <= 2, E_ARG,
               "spread_tiled must be 0 (atomics), 1 (largest tiles) or 2 (smallest tiles)");
  return base;
}

static void tuning_to(const Tuning &tu, int32_t *t, int32_t n) {
  const int32_t v[6] = {tu.lds_elems, tu.pair_waves, tu.spread_tiled, tu.xcd_remap, tu.cell_order, tu.persistent};
  for (int k = 0; k < n && k < 6; ++k) t[k] = v[k];
}

extern "C" int orac_hip_batch_set_tuning(orac_hip_ctx *ctx, int32_t batch, const int32_t *tuning, int32_t n) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(tuning || n == 0, E_ARG, "null tuning");
    Batch &b = get_batch(ctx, batch);
    b.tu = tuning_from(tuning, n, b.tu);
    batch_upload(ctx, b);
  });
}

extern "C" int orac_hip_batch_get_tuning(orac_hip_ctx *ctx, int32_t batch, int32_t *tuning, int32_t n) {
  return guarded([&] {
    ORAC_REQUIRE(tuning && n >= 0, E_ARG, "null tuning");
    tuning_to(get_batch(ctx, batch).tu, tuning, n);
  });
}

// Autotuner: times the batch's step (list rebuild included when asked) for
// every candidate of FFT tile size, pair-kernel occupancy, spread path and
// XCD mapping, one parameter at a time from the current choice (coordinate
// descent, two sweeps), and keeps the fastest.  Candidates never change
// forces; the FFT tile size changes energies only in their last bits.
extern "C" int orac_hip_batch_autotune(orac_hip_ctx *ctx, int32_t batch, int32_t reps, int32_t rebuild_list,
                                       double *best_ms) {
  return guarded([&] {
    set_device(ctx);
    Batch &b = get_batch(ctx, batch);
    const int r = std::max(1, (int)reps);
    const bool rebuild = rebuild_list != 0;
    const bool prof = b.profile;
    b.profile = false;
    Tuning best = b.tu;
    batch_upload(ctx, b);
    double tbest = batch_time(ctx, b, r, rebuild);
    b.tune_log.assign(1, {best, tbest});
    const std::vector<int> cands[6] = {{256, 512, 1024}, {2, 3, 4}, {1, 2, 0}, {1, 0}, {0, 1}, {0, 1}};
    for (int sweep = 0; sweep < 2; ++sweep) {
      for (int param = 0; param < 6; ++param) {
        const std::vector<int> &cand = cands[param];
        for (int v : cand) {
          Tuning t = best;
          int *slots[6] = {&t.lds_elems, &t.pair_waves, &t.spread_tiled, &t.xcd_remap, &t.cell_order,
                           &t.persistent};
          int &slot = *slots[param];
          if (slot == v) continue;
          slot = v;
          b.tu = t;
          batch_upload(ctx, b);
          const double tm = batch_time(ctx, b, r, rebuild);
          b.tune_log.push_back({t, tm});
          if (tm < tbest) {
            tbest = tm;
            best = t;
          }
        }
      }
    }
    b.tu = best;
    batch_upload(ctx, b);
    b.profile = prof;
    if (best_ms) *best_ms = tbest;
  });
}

extern "C" int orac_hip_batch_tune_log(orac_hip_ctx *ctx, int32_t batch, int32_t *tunings, double *ms, int32_t cap,
                                       int32_t *n) {
  return guarded([&] {
    ORAC_REQUIRE(n && cap >= 0 && (cap == 0 || (tunings && ms)), E_ARG, "null argument");
    const Batch &b = get_batch(ctx, batch);
    *n = (int32_t)b.tune_log.size();
    for (int k = 0; k < cap && k < *n; ++k) {
      tuning_to(b.tune_log[k].first, tunings + 6 * (size_t)k, 6);
      ms[k] = b.tune_log[k].second;
    }
  });
}

// ================================================================ C ABI: two nonbonded shells

extern "C" int orac_hip_system_set_shells(orac_hip_ctx *ctx, int32_t sys, double rcutm, double rtolm) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    hipStream_t st = ctx->stream;
    PairDesc &p = S.pair;
    if (rcutm <= 0.0) {  // back to one shell: rinn = 0 (mts_forces.f:88-93)
      S.rcutm = S.rtolm = 0.0;
      p.r2inn = p.rinn0 = p.r2inn0 = p.arsinn1 = p.arsinn2 = 0.0;
      p.lskip_ewald = 0;
      S.list.r2m = 0.0;
    } else {
      ORAC_REQUIRE(rtolm >= 0.0 && rcutm + rtolm < S.rcut, E_ARG, "the inner shell must end before the cut-off");
      ORAC_REQUIRE(!S.alch && p.rematom == nullptr && !p.half, E_UNSUPPORTED,
                   "two shells are built for plain (not alchemical, REM or half-list) systems");
      S.rcutm = rcutm;
      S.rtolm = rtolm;
      // outer shell: rinn = rcutm switched in over rtolm (force.f:145-166); its
      // real-space Ewald term is skipped where erfc(alpha rcutm) < 1e-3
      // (decide_force, mts_forces.f:110-115)
      p.r2inn = rcutm * rcutm;
      p.rinn0 = rcutm + rtolm;
      p.r2inn0 = p.rinn0 * p.rinn0;
      p.arsinn1 = p.rinn0 - 3.0 * rcutm;
      p.arsinn2 = rtolm * rtolm * rtolm;
      int32_t skip = 0;
      ORAC_REQUIRE(orac_hip_decide_force(rcutm, S.alpha, &skip) == 0, E_STATE, "decide_force failed");
      p.lskip_ewald = skip;
      // inner shell: rinn = 0, rout = rcutm switched out over rtolm
      PairDesc &m = S.pair_m;
      m = p;
      m.r2inn = m.rinn0 = m.r2inn0 = m.arsinn1 = m.arsinn2 = 0.0;
      m.r2out = rcutm * rcutm;
      m.rout0 = rcutm + rtolm;
      m.r2out0 = m.rout0 * m.rout0;
      m.arsout1 = m.rout0 - 3.0 * rcutm;
      m.arsout2 = rtolm * rtolm * rtolm;
      m.lskip_ewald = 0;
      if (!S.fmx) {
        S.fmx = S.B.get<double>("fmx", S.natoms);
        S.fmy = S.B.get<double>("fmy", S.natoms);
        S.fmz = S.B.get<double>("fmz", S.natoms);
        S.out_m = S.B.get<double>("out_m", 16);
        for (double *q : {S.fmx, S.fmy, S.fmz}) launch_zero_f64(q, S.natoms, st);
        launch_zero_f64(S.out_m, 16, st);
      }
      m.fx = S.fmx;
      m.fy = S.fmy;
      m.fz = S.fmz;
      m.out = S.out_m;
      m.part = S.B.get<double>("ppart_m", (size_t)m.nblk * 16);
      // the inner shell walks an inner list: the entries of the full list within
      // rcutm + rtolm + the skin, filtered whenever the full list is rebuilt
      ORAC_REQUIRE(S.list.slab > 0 && !S.list.split, E_STATE, "two shells need the slab list layout");
      const double rm = rcutm + rtolm + S.rnei;
      S.list.r2m = rm * rm;
      S.list.out_m = S.B.get<int>("nbidx_m", S.list_cap);
      S.list.count_m = S.B.get<int>("lcount_m", S.ngrp);
      m.nb_idx = S.list.out_m;
      m.nb_cnt = S.list.count_m;
      m.nb_slab = S.list.slab;
      S.have_list = false;  // the next evaluation builds both lists
    }
    if (S.md_on) {
      S.md.fmx = S.rcutm > 0.0 ? S.fmx : nullptr;
      S.md.fmy = S.rcutm > 0.0 ? S.fmy : nullptr;
      S.md.fmz = S.rcutm > 0.0 ? S.fmz : nullptr;
      S.md_d = S.B.put("md_desc", &S.md, 1, st, false);
      S.md_primed = false;
      engine(ctx).md_gen++;
    }
    ORAC_HIP(hipStreamSynchronize(st));
    S.gen++;  // every batch of the system uploads its descriptors again
  });
}

// ================================================================ C ABI: dynamics

extern "C" int orac_hip_md_setup(orac_hip_ctx *ctx, int32_t sys, int32_t nrigid, const int32_t *rigid_first,
                                 const double *rigid_len, int32_t nterm, const int32_t *terms, const double *term_k,
                                 const double *term_r0) {
  return guarded([&] {
    set_device(ctx);
    md_setup(ctx, get_sys(ctx, sys), nrigid, rigid_first, rigid_len, nterm, terms, term_k, term_r0);
  });
}

extern "C" int orac_hip_md_set_state(orac_hip_ctx *ctx, int32_t sys, const double *x, const double *y,
                                     const double *z, const double *vx, const double *vy, const double *vz) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_REQUIRE(S.md_on, E_STATE, "dynamics is not set up for this system");
    hipStream_t st = ctx->stream;
    const int n = S.natoms;
    const size_t nb = (size_t)n * sizeof(double);
    if (x && y && z) {
      std::vector<double> xp(n), yp(n), zp(n);
      const double *o = S.oc;
      for (int a = 0; a < n; ++a) {
        xp[a] = o[0] * x[a] + o[3] * y[a] + o[6] * z[a];
        yp[a] = o[1] * x[a] + o[4] * y[a] + o[7] * z[a];
        zp[a] = o[2] * x[a] + o[5] * y[a] + o[8] * z[a];
      }
      ORAC_HIP(hipMemcpyAsync(S.md.x, x, nb, hipMemcpyHostToDevice, st));
      ORAC_HIP(hipMemcpyAsync(S.md.y, y, nb, hipMemcpyHostToDevice, st));
      ORAC_HIP(hipMemcpyAsync(S.md.z, z, nb, hipMemcpyHostToDevice, st));
      ORAC_HIP(hipMemcpyAsync(S.xp, xp.data(), nb, hipMemcpyHostToDevice, st));
      ORAC_HIP(hipMemcpyAsync(S.yp, yp.data(), nb, hipMemcpyHostToDevice, st));
      ORAC_HIP(hipMemcpyAsync(S.zp, zp.data(), nb, hipMemcpyHostToDevice, st));
      S.have_list = false;
      S.md_primed = false;
    }
    const bool moved = x && y && z;
    if (vx && vy && vz) {
      ORAC_HIP(hipMemcpyAsync(S.md.vx, vx, nb, hipMemcpyHostToDevice, st));
      ORAC_HIP(hipMemcpyAsync(S.md.vy, vy, nb, hipMemcpyHostToDevice, st));
      ORAC_HIP(hipMemcpyAsync(S.md.vz, vz, nb, hipMemcpyHostToDevice, st));
      // velocities onto the constraint surface of the rigid units (RATTLE)
      launch_md_units(S.md_d, &S.md, 1, 0.0, 0.0, 0.0, MD_RATTLE, st);
    }
    ORAC_HIP(hipStreamSynchronize(st));
    if (moved) size_list(ctx, sys);
  });
}

extern "C" int orac_hip_md_get_state(orac_hip_ctx *ctx, int32_t sys, double *x, double *y, double *z, double *vx,
                                     double *vy, double *vz) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_REQUIRE(S.md_on, E_STATE, "dynamics is not set up for this system");
    hipStream_t st = ctx->stream;
    const size_t nb = (size_t)S.natoms * sizeof(double);
    double *dst[6] = {x, y, z, vx, vy, vz};
    const double *src[6] = {S.md.x, S.md.y, S.md.z, S.md.vx, S.md.vy, S.md.vz};
    for (int k = 0; k < 6; ++k)
      if (dst[k]) ORAC_HIP(hipMemcpyAsync(dst[k], src[k], nb, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
  });
}

extern "C" int orac_hip_batch_md_run(orac_hip_ctx *ctx, int32_t batch, int32_t nsteps, double dt, int32_t ninner,
                                     int32_t list_every, int32_t use_graph) {
  return guarded([&] {
    set_device(ctx);
    md_run(ctx, get_batch(ctx, batch), nsteps, dt, 0, ninner, list_every, use_graph != 0);
  });
}

extern "C" int orac_hip_batch_md_run_shells(orac_hip_ctx *ctx, int32_t batch, int32_t nsteps, double dt, int32_t nmid,
                                            int32_t ninner, int32_t list_every, int32_t use_graph) {
  return guarded([&] {
    set_device(ctx);
    ORAC_REQUIRE(nmid >= 1, E_ARG, "nmid must be at least 1");
    md_run(ctx, get_batch(ctx, batch), nsteps, dt, nmid, ninner, list_every, use_graph != 0);
  });
}

extern "C" int orac_hip_batch_md_relax(orac_hip_ctx *ctx, int32_t batch, int32_t niter, double hstep, double dmax,
                                       int32_t list_every) {
  return guarded([&] {
    set_device(ctx);
    Batch &b = get_batch(ctx, batch);
    ORAC_REQUIRE(niter >= 0 && hstep > 0.0 && dmax > 0.0 && list_every >= 1, E_ARG, "bad relaxation parameters");
    md_upload(ctx, b, b.md_every > 0 ? b.md_every : 1);
    const int n = (int)b.sys.size();
    for (int it = 0; it < niter; ++it) {
      if (it % list_every == 0 && it > 0) {
        // the positions move far while relaxing: a list that outgrew its slab
        // is resized at the current positions before the next rebuild
        for (int id : list_overflows(ctx, b)) size_list(ctx, id);
        md_upload(ctx, b, b.md_every > 0 ? b.md_every : 1);
      }
      launch_md_fast_forces(b.dm, b.hm.data(), n, ctx->stream);
      batch_eval(ctx, b, it % list_every == 0, false);
      if (b.shells) batch_enqueue_mid(ctx, b);
      launch_md_relax(b.dm, b.hm.data(), n, hstep, dmax, ctx->stream);
    }
    ORAC_HIP(hipGetLastError());
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
    for (int id : list_overflows(ctx, b)) size_list(ctx, id);
    for (System *S : b.sys) S->md_primed = false;
  });
}

extern "C" int orac_hip_batch_md_scale_velocities(orac_hip_ctx *ctx, int32_t batch, const double *scale) {
  return guarded([&] {
    set_device(ctx);
    Batch &b = get_batch(ctx, batch);
    ORAC_REQUIRE(scale, E_ARG, "null scale");
    md_upload(ctx, b, b.md_every > 0 ? b.md_every : 1);
    double *d = b.B.put("md_scale", scale, b.sys.size(), ctx->stream, false);
    launch_md_scale_v(b.dm, b.hm.data(), (int)b.sys.size(), d, ctx->stream);
    ORAC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int orac_hip_system_md_energies(orac_hip_ctx *ctx, int32_t sys, double *out, int32_t *fail) {
  return guarded([&] {
    set_device(ctx);
    System &S = get_sys(ctx, sys);
    ORAC_REQUIRE(S.md_on, E_STATE, "dynamics is not set up for this system");
    hipStream_t st = ctx->stream;
    launch_md_fast_forces(S.md_d, &S.md, 1, st);
    launch_md_energy(S.md_d, 1, st);
    double o[5];
    int f[2];
    ORAC_HIP(hipMemcpyAsync(o, S.md.out, sizeof o, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipMemcpyAsync(f, S.md.fail, sizeof f, hipMemcpyDeviceToHost, st));
    ORAC_HIP(hipStreamSynchronize(st));
    if (out)
      for (int k = 0; k < 5; ++k) out[k] = o[k];
    if (fail) *fail = f[0];
  });
}
