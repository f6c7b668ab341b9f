// C-ABI of the library: per-cell drop-ins for the reference solver plugin
// surface and the device-resident grid.  See include/exahype_amd.h.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "aderdg_kernels.h"
#include "grid_kernels.h"
#include "predictor_tu.h"
#include "grid_view.h"
#include "host_util.h"
#include "solver_internal.h"
#include "dg_tables_data.h"


static thread_local std::string g_err;

void exa_set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

// Order dispatch: instantiates a lambda-like body for the supported orders.
#define EXA_ORDER_SWITCH(N, BODY)                 \
  switch (N) {                                    \
    case 1: { constexpr int NN = 1; BODY; } break; \
    case 2: { constexpr int NN = 2; BODY; } break; \
    case 3: { constexpr int NN = 3; BODY; } break; \
    case 4: { constexpr int NN = 4; BODY; } break; \
    case 5: { constexpr int NN = 5; BODY; } break; \
    case 6: { constexpr int NN = 6; BODY; } break; \
    case 7: { constexpr int NN = 7; BODY; } break; \
    case 8: { constexpr int NN = 8; BODY; } break; \
    case 9: { constexpr int NN = 9; BODY; } break; \
    default:                                      \
      exa_set_error("order %d not supported (1..9)", N); \
      return EXA_ERR_UNSUPPORTED;                 \
  }

extern "C" const char* exa_last_error(void) { return g_err.c_str(); }
extern "C" const char* exa_version(void) { return "exahype_amd 0.1 gfx950"; }

// ---------------------------------------------------------------------------
// device initialisation: tables to constant memory, large-LDS attributes
static std::mutex g_init_mu;
static unsigned long long g_init_mask = 0;
static DGTables g_host_tab[EXA_MAX_ORDER + 1];  // the DG tables uploaded by exa_init

// Predictor kernels live in the per-order units pred_n<N>.hip.
template <int N>
static PredFn predictor_kernel_src(bool debug) {
  switch (N) {
    case 1: return exa_pred_kernel_src_1(debug);
    case 2: return exa_pred_kernel_src_2(debug);
    case 3: return exa_pred_kernel_src_3(debug);
    case 4: return exa_pred_kernel_src_4(debug);
    case 5: return exa_pred_kernel_src_5(debug);
    case 6: return exa_pred_kernel_src_6(debug);
    case 7: return exa_pred_kernel_src_7(debug);
    default: return nullptr;
  }
}

template <int N>
static PredFn predictor_kernel(bool debug, int prec, int wpe, int form, bool uskip = false) {
  switch (N) {
    case 1: return exa_pred_kernel_1(debug, prec, wpe, form, uskip);
    case 2: return exa_pred_kernel_2(debug, prec, wpe, form, uskip);
    case 3: return exa_pred_kernel_3(debug, prec, wpe, form, uskip);
    case 4: return exa_pred_kernel_4(debug, prec, wpe, form, uskip);
    case 5: return exa_pred_kernel_5(debug, prec, wpe, form, uskip);
    case 6: return exa_pred_kernel_6(debug, prec, wpe, form, uskip);
    case 7: return exa_pred_kernel_7(debug, prec, wpe, form, uskip);
    default: return nullptr;  // orders 8, 9: exa_pred_staged_kernel
  }
}

// Staging slots of the staged predictor: one per resident workgroup.
static int staged_slots(int N, int device, long long ncells, int* slots) {
  int ncu = 0, per_cu = 0;
  EXA_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
  const int nt = round_up64((N + 1) * (N + 1) * (N + 1));
  EXA_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(exa_pred_staged_kernel(N, false)), nt,
      (size_t)exa_pred_staged_lds_bytes(N)));
  long long n = (long long)ncu * (per_cu > 0 ? per_cu : 1);
  if (n > ncells) n = ncells;
  *slots = (int)(n > 0 ? n : 1);
  return EXA_OK;
}
static size_t staged_slot_bytes(int N) {
  const size_t B = (size_t)N + 1;
  return 2 * B * B * B * B * exa::V * sizeof(double);
}

// LDS of the stand-alone volume integral: sQh slot + sFh, 4 V B^3 doubles
template <int N>
constexpr int volume_integral_lds_bytes() {
  return 4 * exa::V * exa::Ord<N>::B3 * (int)sizeof(double);
}
template <int N>
static hipError_t set_lds_attrs() {
  const int bytes = volume_integral_lds_bytes<N>();
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&exa::k_volume_integral_ref<N>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

extern "C" int exa_init(int device) {
  int n = 0;
  EXA_CHECK(hipGetDeviceCount(&n));
  EXA_ARG(device >= 0 && device < n && device < 64, "device index");
  EXA_CHECK(hipSetDevice(device));
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_init_mask & (1ull << device)) return EXA_OK;
  hipDeviceProp_t prop;
  EXA_CHECK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    exa_set_error("device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
    return EXA_ERR_UNSUPPORTED;
  }
  DGTables* host = g_host_tab;
  std::memset(host, 0, sizeof(g_host_tab));
  for (int N = 0; N <= EXA_MAX_ORDER; ++N) {
    for (int i = 0; i < EXA_NB; ++i) {
      host[N].w[i] = EXA_TAB_gl_weights[N][i];
      host[N].xi[i] = EXA_TAB_gl_nodes[N][i];
      host[N].F0[i] = EXA_TAB_F0[N][i];
      host[N].FL[i] = EXA_TAB_FLCoeff[N][i];
      host[N].FR[i] = EXA_TAB_FRCoeff[N][i];
      for (int j = 0; j < EXA_NB; ++j) {
        host[N].Kxi[i][j] = EXA_TAB_Kxi[N][i][j];
        host[N].iK1[i][j] = EXA_TAB_iK1[N][i][j];
        host[N].dudx[i][j] = EXA_TAB_dudx[N][i][j];
        host[N].eq[i][j] = EXA_TAB_equidistantGridProjector1d[N][i][j];
      }
    }
  }
  EXA_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_tab), host, sizeof(g_host_tab)));
  static DGTablesF hostf[EXA_MAX_ORDER + 1];
  std::memset(hostf, 0, sizeof(hostf));
  for (int N = 0; N <= EXA_MAX_ORDER; ++N)
    for (int i = 0; i < EXA_NB; ++i) {
      hostf[N].w[i] = (float)host[N].w[i];
      hostf[N].F0[i] = (float)host[N].F0[i];
      for (int j = 0; j < EXA_NB; ++j) {
        hostf[N].Kxi[i][j] = (float)host[N].Kxi[i][j];
        hostf[N].iK1[i][j] = (float)host[N].iK1[i][j];
      }
    }
  EXA_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_tabf), hostf, sizeof(hostf)));
  EXA_CHECK(exa_pred_init_1(host, hostf));
  EXA_CHECK(exa_pred_init_2(host, hostf));
  EXA_CHECK(exa_pred_init_3(host, hostf));
  EXA_CHECK(exa_pred_init_4(host, hostf));
  EXA_CHECK(exa_pred_init_5(host, hostf));
  EXA_CHECK(exa_pred_init_6(host, hostf));
  EXA_CHECK(exa_pred_init_7(host, hostf));
  EXA_CHECK(exa_pred_staged_init(host, hostf));
  EXA_CHECK(exa_pde_init(host));
  if (exa_pde_user_init_all(host) != 0) {
    exa_set_error("uploading the DG tables of a registered PDE failed on device %d", device);
    return EXA_ERR_HIP;
  }
  EXA_TRY(exa_amr_noalloc_setup(device));
  EXA_CHECK(exa_amr_grid_init(host));
  EXA_CHECK(set_lds_attrs<1>());
  EXA_CHECK(set_lds_attrs<2>());
  EXA_CHECK(set_lds_attrs<3>());
  EXA_CHECK(set_lds_attrs<4>());
  EXA_CHECK(set_lds_attrs<5>());
  EXA_CHECK(set_lds_attrs<6>());
  EXA_CHECK(set_lds_attrs<7>());
  EXA_CHECK(set_lds_attrs<8>());
  EXA_CHECK(set_lds_attrs<9>());
  g_init_mask |= (1ull << device);
  return EXA_OK;
}

extern "C" int exa_pde_register(const char* name, const exa_pde_kernels* k, int* pde) {
  EXA_ARG(name && name[0] && k && pde, "null pointer or empty name");
  EXA_ARG(std::strlen(name) < 64, "PDE names are shorter than 64 characters");
  EXA_ARG(k->version == EXA_PDE_KERNELS_VERSION, "exa_pde_kernels of another version: rebuild against this header");
  EXA_ARG(k->init && k->max_order >= 1 && k->max_order <= 7, "kernel table without init or with orders beyond 1..7");
  for (int n = 1; n <= k->max_order; ++n)
    EXA_ARG(k->predictor[n][0] && k->predictor[n][1] && k->riemann[n] && k->riemann_list[n] && k->update_dt[n] &&
                k->update_full[n] && k->initial_conditions[n] && k->boundary_faces[n] && k->boundary_ref[n],
            "kernel table with missing entries");
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  const int slot = exa_pde_user_add(name, k);
  EXA_ARG(slot >= 0, "every PDE slot is taken (EXA_PDE_USER_SLOTS)");
  const int id = EXA_PDE_USER_FIRST + slot;
  std::lock_guard<std::mutex> lk(g_init_mu);
  for (int d = 0; d < 64; ++d) {
    if (!(g_init_mask & (1ull << d))) continue;
    EXA_CHECK(hipSetDevice(d));
    if (exa_pde_user_init(id, g_host_tab) != 0) {
      (void)hipSetDevice(dev);
      exa_set_error("uploading the DG tables of PDE '%s' failed on device %d", name, d);
      return EXA_ERR_HIP;
    }
  }
  EXA_CHECK(hipSetDevice(dev));
  *pde = id;
  return EXA_OK;
}

extern "C" int exa_pde_lookup(const char* name, int* pde) {
  EXA_ARG(name && pde, "null pointer");
  const int slot = exa_pde_user_find(name);
  EXA_ARG(slot >= 0, "no PDE of that name is registered");
  *pde = EXA_PDE_USER_FIRST + slot;
  return EXA_OK;
}

extern "C" int exa_device_synchronize(void) {
  EXA_CHECK(hipDeviceSynchronize());
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// solver handle
struct exa_solver {
  exa_aderdg_config cfg;
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  // scratch for the per-cell entry points
  DevBuf luh, lduh, qf, ff, iters, dt, dbg, misc, stage;
  DevBuf psi;  // point forces of one cell (linear predictor with a point source)
  // deferred per-cell work (exa_aderdg_batch_*): host staging until the flush
  struct PredItem { double *lQhbnd, *lFhbnd, *lduh; };
  struct RiemItem { double *FL, *FR; };
  struct UpdItem { double *luh, *lduh; };
  std::vector<PredItem> bp;
  std::vector<double> bp_luh, bp_par;  // SoA state per cell; reference-layout copy when npar > 0
  std::vector<exa::InstParams> bp_inst;
  std::vector<RiemItem> br;
  std::vector<double> br_q, br_f;
  std::vector<int> br_dir;
  std::vector<double> br_dt;
  std::vector<UpdItem> bu;
  std::vector<double> bu_luh, bu_lduh, bu_ff;
  std::vector<exa::InstParams> bu_inst;
  DevBuf b_luh, b_lduh, b_qf, b_ff, b_inst, b_cinst, b_iters, b_stage, b_dirs;
};

// Contraction form of the predictor's Picard sweeps: EXA_CONTRACT_PENCIL
// (two barriers per time slice) or EXA_CONTRACT_NODE (one barrier per slice).
// Predictor variants chosen by exa_autotune, per device, order and precision
// (0: not tuned, the measured defaults below apply).
struct TunedVariant {
  int wpe = 0;
  int contraction = 0;
};
static TunedVariant g_tuned[64][EXA_MAX_ORDER + 1][4];

// The staged predictor: orders 8, 9, chosen explicitly, or the measured
// default of the FP64 predictor at N = 7 (81^3 cells: 222 ms per launch
// staged vs 339 ms register-resident, whose sweeps spill at that order).
static bool predictor_staged(const exa_aderdg_config& c, int dev) {
  if (exa_pde_generic(c.pde)) return false;
  if (c.order >= 8 || c.contraction == EXA_CONTRACT_STAGED) return true;
  return c.order == 7 && c.precision == EXA_PREC_FP64 && c.contraction == EXA_CONTRACT_DEFAULT &&
         g_tuned[dev & 63][7][EXA_PREC_FP64].contraction == 0;  // exa_autotune's choice wins
}

// Launch extras of a single-grid predictor launch: the gravity source of
// EXA_PDE_EULER_GRAVITY (0 for the plain Euler equations).
static bool pde_has_source_or_ncp(int pde) {
  return pde == EXA_PDE_EULER_GRAVITY || pde == EXA_PDE_ADVECTION_NCP || pde == EXA_PDE_EULER_NCP_TEST;
}

static exa::Ens launch_extras(const exa_aderdg_config& c) {
  exa::Ens e{};
  e.grav = c.pde == EXA_PDE_EULER_GRAVITY ? c.gravity : 0.0;
  for (int d = 0; d < 3; ++d) e.adv[d] = c.velocity[d];
  if (c.point_source) {  // the source's parameters; callers choose how the predictor gets it
    for (int d = 0; d < 3; ++d) e.ps.x0[d] = c.source_x0[d];
    for (int v = 0; v < 5; ++v) e.ps.amp[v] = c.source_amp[v];
    e.ps.f0 = c.source_f0;
    e.ps.t0 = c.source_t0;
  }
  return e;
}

static int predictor_form(const exa_aderdg_config& c, int N, int dev) {
  switch (c.contraction) {
    case EXA_CONTRACT_PENCIL: return exa::FORM_PENCIL;
    case EXA_CONTRACT_NODE: return exa::FORM_NODE;
    case EXA_CONTRACT_PENCIL_PAIR: return exa::FORM_PENCIL_PAIR;
    default: break;
  }
  const int tc = g_tuned[dev & 63][N][c.precision & 3].contraction;
  if (tc) return tc == EXA_CONTRACT_NODE ? exa::FORM_NODE
                                         : (tc == EXA_CONTRACT_PENCIL_PAIR ? exa::FORM_PENCIL_PAIR : exa::FORM_PENCIL);
  // measured at 81^3 cells, FP64 predictor ms (pencil / node / pencil pairs):
  // N=3 6.73 / 7.15 / 6.91, N=5 42.1 / 67.4 / 38.1, N=7 355 / 346 / 341
  return N >= 5 ? exa::FORM_PENCIL_PAIR : exa::FORM_PENCIL;
}

extern "C" int exa_aderdg_create(const exa_aderdg_config* cfg, exa_solver** out) {
  EXA_ARG(cfg && out, "null pointer");
  EXA_ARG(cfg->order >= 1 && cfg->order <= 9, "order must be 1..9");
  EXA_ARG(cfg->order <= 7 || cfg->precision <= EXA_PREC_FP32,
          "orders 8 and 9 run the staged predictor: FP64 or FP32 sweeps with FP64 correction");
  EXA_ARG(cfg->contraction != EXA_CONTRACT_STAGED ||
              (cfg->precision <= EXA_PREC_FP32 && cfg->order >= EXA_STAGED_MIN_ORDER),
          "the staged predictor is FP64 or FP32 + FP64 and built for orders 5..9");
  EXA_ARG((cfg->pde == EXA_PDE_EULER || cfg->pde == EXA_PDE_EULER_GRAVITY || exa_pde_generic(cfg->pde)) &&
              cfg->nvar == 5,
          "the compiled PDEs are 3D Euler (optionally with gravity or the test non-conservative product) "
          "and linear advection (flux or non-conservative form), all with 5 variables");
  EXA_ARG(!exa_pde_generic(cfg->pde) || (cfg->precision == EXA_PREC_FP64 && (cfg->linear || cfg->order <= 7) &&
                                         cfg->contraction == EXA_CONTRACT_DEFAULT && cfg->npar == 0),
          "the advection and non-conservative-product PDEs run the FP64 predictor at orders 1..7 "
          "(1..9 with the linear kernels; default contraction, no parameters)");
  EXA_ARG(!cfg->linear || exa_pde_linear_supported(cfg->pde),
          "the linear kernels need a PDE with a coefficient matrix: EXA_PDE_ADVECTION_NCP or "
          "EXA_PDE_EULER_NCP_TEST");
  EXA_ARG(!cfg->point_source || cfg->linear,
          "point sources belong to the linear kernels (the usePointSource arm of spaceTimePredictorLinear)");
  EXA_ARG(cfg->pde != EXA_PDE_EULER_GRAVITY || cfg->precision == EXA_PREC_FP64,
          "the gravity source runs in the FP64 predictor only");
  EXA_ARG(cfg->npar >= 0 && cfg->npar <= 3, "npar must be 0..3");
  EXA_ARG(cfg->cfl > 0.0 && cfg->gamma > 1.0, "cfl/gamma");
  EXA_ARG(cfg->precision >= EXA_PREC_FP64 && cfg->precision <= EXA_PREC_BF16_MFMA, "precision");
  EXA_ARG(cfg->precision < EXA_PREC_FP32_MFMA || cfg->order <= 7, "MFMA sweeps need order <= 7");
  EXA_ARG(cfg->waves_per_simd >= 0 && cfg->waves_per_simd <= 2, "waves_per_simd must be 0, 1 or 2");
  EXA_ARG(cfg->contraction >= 0 && cfg->contraction <= 4, "contraction must be 0..4");
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  exa_solver* s = new exa_solver();
  s->cfg = *cfg;
  if (s->cfg.fp64_sweeps < 1) s->cfg.fp64_sweeps = 1;
  s->device = dev;
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    exa_set_error("hipStreamCreate: %s", hipGetErrorString(e));
    return EXA_ERR_HIP;
  }
  *out = s;
  return EXA_OK;
}

extern "C" int exa_aderdg_destroy(exa_solver* s) {
  if (!s) return EXA_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return EXA_OK;
}

// ---- layout conversion helpers (reference AoS <-> device SoA) ----
static void ref_to_soa_cell(const double* ref, double* soa, int nnodes, int nd) {
  for (int n = 0; n < nnodes; ++n)
    for (int v = 0; v < exa::V; ++v) soa[v * nnodes + n] = ref[n * nd + v];
}
static void soa_to_ref_cell(const double* soa, double* ref, int nnodes, int nd) {
  for (int n = 0; n < nnodes; ++n)
    for (int v = 0; v < exa::V; ++v) ref[n * nd + v] = soa[v * nnodes + n];
}
// faces: reference [face][a][b][nd]  <->  device [face][v][ab]
static void ref_to_soa_faces(const double* ref, double* soa, int nfaces, int B2, int nd) {
  for (int f = 0; f < nfaces; ++f) ref_to_soa_cell(ref + (size_t)f * B2 * nd, soa + (size_t)f * exa::V * B2, B2, nd);
}
static void soa_to_ref_faces(const double* soa, double* ref, int nfaces, int B2, int nd) {
  for (int f = 0; f < nfaces; ++f) soa_to_ref_cell(soa + (size_t)f * exa::V * B2, ref + (size_t)f * B2 * nd, B2, nd);
}

static int sync_stream(exa_solver* s) {
  EXA_CHECK(hipStreamSynchronize(s->stream));
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

// Material parameters are constant in time and carried as the reference does:
// the trivial guess copies them, time averaging and face extrapolation weight
// them.  Writes their entries of lQhbnd; parh = their time averages [node][P].
static void param_predictor(int N, int nd, const double* luh, double* lQhbnd, std::vector<double>& parh) {
  const int B = N + 1, B3 = B * B * B, V = exa::V, P = nd - V;
  parh.assign((size_t)B3 * P, 0.0);
  const double* w = EXA_TAB_gl_weights[N];
  for (int n = 0; n < B3; ++n)
    for (int p = 0; p < P; ++p) {
      double a = 0.0;
      for (int t = 0; t < B; ++t) a += luh[n * nd + V + p] * w[t];
      parh[n * P + p] = a;
    }
  const double* FLc = EXA_TAB_FLCoeff[N];
  const double* FRc = EXA_TAB_FRCoeff[N];
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < B; ++a)
      for (int b = 0; b < B; ++b)
        for (int p = 0; p < P; ++p) {
          double qL = 0.0, qR = 0.0;
          for (int l = 0; l < B; ++l) {
            const int z = d == 2 ? l : a, y = d == 0 ? b : (d == 1 ? l : a), x = d == 0 ? l : b;
            const double v = parh[((z * B + y) * B + x) * P + p];
            qL += v * FLc[l];
            qR += v * FRc[l];
          }
          lQhbnd[(((2 * d) * B + a) * B + b) * nd + V + p] = qL;
          lQhbnd[(((2 * d + 1) * B + a) * B + b) * nd + V + p] = qR;
        }
}

// spaceTimePredictorLinear of one cell (+ its volume integral, not returned).
static int linear_predictor_cell(exa_solver* s, double* lQhbnd, double* lFhbnd, double* lQi, double* lFi,
                                 double* lQhi, double* lFhi, const double* luh, const double dx[3], double dt,
                                 const double* point_forces, int* picard_iterations) {
  const int N = s->cfg.order, B = N + 1, B2 = B * B, B3 = B2 * B, B4 = B2 * B2, V = exa::V;
  std::vector<double> h(std::max<size_t>((size_t)V * B3, (size_t)6 * V * B2));
  ref_to_soa_cell(luh, h.data(), B3, V);
  EXA_TRY(s->luh.ensure(sizeof(double) * V * B3));
  EXA_TRY(s->lduh.ensure(sizeof(double) * V * B3));
  EXA_TRY(s->qf.ensure(sizeof(double) * 6 * V * B2));
  EXA_TRY(s->ff.ensure(sizeof(double) * 6 * V * B2));
  EXA_TRY(s->iters.ensure(sizeof(int) * 4));
  EXA_TRY(s->dt.ensure(sizeof(double) * 4));
  const size_t nQi = (size_t)(B + 1) * B3 * V, nFi = (size_t)B4 * 3 * V, nQhi = (size_t)B3 * V,
               nFhi = (size_t)4 * B3 * V;
  EXA_TRY(s->dbg.ensure(sizeof(double) * (nQi + nFi + nQhi + nFhi)));
  exa::PredictorDebug dbg{s->dbg.as<double>(), s->dbg.as<double>() + nQi, s->dbg.as<double>() + nQi + nFi,
                          s->dbg.as<double>() + nQi + nFi + nQhi};
  EXA_CHECK(hipMemcpyAsync(s->luh.p, h.data(), sizeof(double) * V * B3, hipMemcpyHostToDevice, s->stream));
  EXA_CHECK(hipMemcpyAsync(s->dt.p, &dt, sizeof(double), hipMemcpyHostToDevice, s->stream));
  exa::Ens ens = launch_extras(s->cfg);
  if (point_forces) {
    const size_t np = (size_t)(B + 1) * B3 * V;
    EXA_TRY(s->psi.ensure(sizeof(double) * np));
    EXA_CHECK(hipMemcpyAsync(s->psi.p, point_forces, sizeof(double) * np, hipMemcpyHostToDevice, s->stream));
    ens.ps.mode = 1;
    ens.ps.psi = s->psi.as<double>();
  }
  hipLaunchKernelGGL(exa_pde_linear_predictor(s->cfg.pde, N, true, point_forces != nullptr), dim3(1),
                     dim3(round_up64(B3)),
                     (size_t)exa_pde_linear_lds_bytes(N), s->stream, s->luh.as<double>(), s->lduh.as<double>(),
                     s->qf.as<double>(), s->ff.as<double>(), s->iters.as<int>(), s->dt.as<double>(), dx[0], dx[1],
                     dx[2], s->cfg.gamma, nullptr, dbg, ens);
  EXA_CHECK(hipGetLastError());
  std::vector<double> hq(6 * V * B2), hf(6 * V * B2), hdbg(nQi + nFi + nQhi + nFhi);
  EXA_CHECK(hipMemcpyAsync(hq.data(), s->qf.p, sizeof(double) * 6 * V * B2, hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipMemcpyAsync(hf.data(), s->ff.p, sizeof(double) * 6 * V * B2, hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipMemcpyAsync(hdbg.data(), s->dbg.p, sizeof(double) * hdbg.size(), hipMemcpyDeviceToHost, s->stream));
  EXA_TRY(sync_stream(s));
  soa_to_ref_faces(hq.data(), lQhbnd, 6, B2, V);
  soa_to_ref_faces(hf.data(), lFhbnd, 6, B2, V);
  if (lQi) std::memcpy(lQi, hdbg.data(), sizeof(double) * nQi);
  if (lFi) std::memcpy(lFi, hdbg.data() + nQi, sizeof(double) * nFi);
  if (lQhi) std::memcpy(lQhi, hdbg.data() + nQi + nFi, sizeof(double) * nQhi);
  if (lFhi) std::memcpy(lFhi, hdbg.data() + nQi + nFi + nQhi, sizeof(double) * nFhi);
  if (picard_iterations) *picard_iterations = 0;
  return EXA_OK;
}

extern "C" int exa_aderdg_point_source(exa_solver* s, double t, double dt, const double centre[3],
                                       const double dx[3], double* point_forces) {
  EXA_ARG(s && centre && dx && point_forces, "null pointer");
  EXA_ARG(s->cfg.point_source, "the solver has no point source (config point_source = 0)");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  const int N = s->cfg.order, B = N + 1, B3 = B * B * B;
  const size_t np = (size_t)(B + 1) * B3 * exa::V;
  EXA_TRY(s->psi.ensure(sizeof(double) * np));
  hipLaunchKernelGGL(exa_pde_point_source(N), dim3(1), dim3(round_up64(B3)), 0, s->stream, t, dt, centre[0],
                     centre[1], centre[2], dx[0], dx[1], dx[2], s->psi.as<double>(), launch_extras(s->cfg));
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipMemcpyAsync(point_forces, s->psi.p, sizeof(double) * np, hipMemcpyDeviceToHost, s->stream));
  return sync_stream(s);
}

static int space_time_predictor(exa_solver* s, double* lQhbnd, double* lFhbnd, double* lQi, double* lFi,
                                double* lQhi, double* lFhi, const double* luh, const double dx[3], double dt,
                                const double* point_forces, int* picard_iterations);

extern "C" int exa_aderdg_space_time_predictor(exa_solver* s, double* lQhbnd, double* lFhbnd, double* lQi,
                                               double* lFi, double* lQhi, double* lFhi, const double* luh,
                                               const double dx[3], double dt, int* picard_iterations) {
  return space_time_predictor(s, lQhbnd, lFhbnd, lQi, lFi, lQhi, lFhi, luh, dx, dt, nullptr, picard_iterations);
}

extern "C" int exa_aderdg_space_time_predictor_ps(exa_solver* s, double* lQhbnd, double* lFhbnd, double* lQi,
                                                  double* lFi, double* lQhi, double* lFhi, const double* luh,
                                                  const double dx[3], double dt, const double* point_forces,
                                                  int* picard_iterations) {
  EXA_ARG(s, "null pointer");
  EXA_ARG(!point_forces || s->cfg.linear,
          "point forces need a linear solver (the nonlinear predictor has no point-source arm)");
  return space_time_predictor(s, lQhbnd, lFhbnd, lQi, lFi, lQhi, lFhi, luh, dx, dt, point_forces,
                              picard_iterations);
}

static int space_time_predictor(exa_solver* s, double* lQhbnd, double* lFhbnd, double* lQi, double* lFi,
                                double* lQhi, double* lFhi, const double* luh, const double dx[3], double dt,
                                const double* point_forces, int* picard_iterations) {
  EXA_ARG(s && lQhbnd && lFhbnd && luh && dx, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->cfg.linear) {
    EXA_CHECK(hipSetDevice(s->device));
    return linear_predictor_cell(s, lQhbnd, lFhbnd, lQi, lFi, lQhi, lFhi, luh, dx, dt, point_forces,
                                 picard_iterations);
  }
  EXA_CHECK(hipSetDevice(s->device));
  const int N = s->cfg.order, B = N + 1, B2 = B * B, B3 = B2 * B, B4 = B2 * B2;
  const int nd = s->cfg.nvar + s->cfg.npar, V = exa::V, P = s->cfg.npar;
  std::vector<double> h(std::max<size_t>((size_t)V * B3, (size_t)6 * V * B2));
  ref_to_soa_cell(luh, h.data(), B3, nd);
  EXA_TRY(s->luh.ensure(sizeof(double) * V * B3));
  EXA_TRY(s->lduh.ensure(sizeof(double) * V * B3));
  EXA_TRY(s->qf.ensure(sizeof(double) * 6 * V * B2));
  EXA_TRY(s->ff.ensure(sizeof(double) * 6 * V * B2));
  EXA_TRY(s->iters.ensure(sizeof(int) * 4));
  EXA_TRY(s->dt.ensure(sizeof(double) * 4));
  const size_t nQi = (size_t)B4 * V, nFi = (size_t)B4 * 4 * V, nQhi = (size_t)B3 * V, nFhi = (size_t)4 * B3 * V;
  EXA_TRY(s->dbg.ensure(sizeof(double) * (nQi + nFi + nQhi + nFhi)));
  exa::PredictorDebug dbg{s->dbg.as<double>(), s->dbg.as<double>() + nQi, s->dbg.as<double>() + nQi + nFi,
                          s->dbg.as<double>() + nQi + nFi + nQhi};
  EXA_CHECK(hipMemcpyAsync(s->luh.p, h.data(), sizeof(double) * V * B3, hipMemcpyHostToDevice, s->stream));
  EXA_CHECK(hipMemcpyAsync(s->dt.p, &dt, sizeof(double), hipMemcpyHostToDevice, s->stream));
  const double gamma = s->cfg.gamma;
  if (exa_pde_generic(s->cfg.pde)) {
    hipLaunchKernelGGL(exa_pde_predictor(s->cfg.pde, N, true), dim3(1), dim3(round_up64(B3)),
                       (size_t)exa_pde_predictor_lds_bytes(N), s->stream, s->luh.as<double>(), s->lduh.as<double>(),
                       s->qf.as<double>(), s->ff.as<double>(), s->iters.as<int>(), s->dt.as<double>(), dx[0], dx[1],
                       dx[2], gamma, nullptr, dbg, launch_extras(s->cfg));
  } else EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    if (predictor_staged(s->cfg, s->device)) {
      EXA_TRY(s->stage.ensure(staged_slot_bytes(NN)));
      hipLaunchKernelGGL(exa_pred_staged_kernel(NN, true, s->cfg.pde == EXA_PDE_EULER_GRAVITY,
                                                s->cfg.precision == EXA_PREC_FP32),
                         dim3(1), dim3(O::NT),
                         (size_t)exa_pred_staged_lds_bytes(NN), s->stream, s->luh.as<double>(), s->lduh.as<double>(),
                         s->qf.as<double>(), s->ff.as<double>(), s->iters.as<int>(), s->dt.as<double>(), dx[0],
                         dx[1], dx[2], gamma, 1LL, s->stage.as<double>(), nullptr, dbg, launch_extras(s->cfg),
                         s->cfg.fp64_sweeps);
    } else {
      const int form = predictor_form(s->cfg, NN, s->device);
      PredFn kern = s->cfg.pde == EXA_PDE_EULER_GRAVITY ? predictor_kernel_src<NN>(true)
                                                          : predictor_kernel<NN>(true, s->cfg.precision, 1, form);
      hipLaunchKernelGGL(kern, dim3(1), dim3(O::NT), O::LDS_DOUBLES * sizeof(double), s->stream, s->luh.as<double>(),
                         s->lduh.as<double>(), s->qf.as<double>(), s->ff.as<double>(), s->iters.as<int>(),
                         s->dt.as<double>(), dx[0], dx[1], dx[2], gamma, s->cfg.fp64_sweeps, nullptr, dbg,
                         launch_extras(s->cfg));
    }
  });
  EXA_CHECK(hipGetLastError());
  std::vector<double> hq(6 * V * B2), hf(6 * V * B2), hdbg(nQi + nFi + nQhi + nFhi);
  int it = 0;
  EXA_CHECK(hipMemcpyAsync(hq.data(), s->qf.p, sizeof(double) * 6 * V * B2, hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipMemcpyAsync(hf.data(), s->ff.p, sizeof(double) * 6 * V * B2, hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipMemcpyAsync(hdbg.data(), s->dbg.p, sizeof(double) * hdbg.size(), hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipMemcpyAsync(&it, s->iters.p, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  EXA_TRY(sync_stream(s));
  soa_to_ref_faces(hq.data(), lQhbnd, 6, B2, nd);
  soa_to_ref_faces(hf.data(), lFhbnd, 6, B2, V);
  const double* gQi = hdbg.data();
  const double* gFi = gQi + nQi;
  const double* gQhi = gFi + nFi;
  const double* gFhi = gQhi + nQhi;
  std::vector<double> parh;
  if (P > 0) param_predictor(N, nd, luh, lQhbnd, parh);
  if (lQi) {
    for (int n = 0; n < B3; ++n)
      for (int t = 0; t < B; ++t) {
        for (int v = 0; v < V; ++v) lQi[(n * B + t) * nd + v] = gQi[(n * B + t) * V + v];
        for (int p = 0; p < P; ++p) lQi[(n * B + t) * nd + V + p] = luh[n * nd + V + p];
      }
  }
  if (lFi) std::memcpy(lFi, gFi, sizeof(double) * nFi);
  if (lQhi) {
    for (int n = 0; n < B3; ++n) {
      for (int v = 0; v < V; ++v) lQhi[n * nd + v] = gQhi[n * V + v];
      for (int p = 0; p < P; ++p) lQhi[n * nd + V + p] = parh[n * P + p];
    }
  }
  if (lFhi) std::memcpy(lFhi, gFhi, sizeof(double) * nFhi);
  if (picard_iterations) *picard_iterations = it;
  return EXA_OK;
}

extern "C" int exa_aderdg_volume_integral(exa_solver* s, double* lduh, const double* lFhi, const double dx[3]) {
  EXA_ARG(s && lduh && lFhi && dx, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  const int N = s->cfg.order, B = N + 1, B3 = B * B * B, V = exa::V;
  EXA_TRY(s->misc.ensure(sizeof(double) * 4 * B3 * V));
  EXA_TRY(s->lduh.ensure(sizeof(double) * B3 * V));
  EXA_CHECK(hipMemcpyAsync(s->misc.p, lFhi, sizeof(double) * (s->cfg.linear ? 3 : 4) * B3 * V,
                           hipMemcpyHostToDevice, s->stream));
  if (s->cfg.linear) {
    hipLaunchKernelGGL(exa_pde_linear_volume(N), dim3(1), dim3(round_up64(B3)), 0, s->stream, s->misc.as<double>(),
                       s->lduh.as<double>());
  } else EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    hipLaunchKernelGGL((exa::k_volume_integral_ref<NN>), dim3(1), dim3(O::NT), (size_t)volume_integral_lds_bytes<NN>(),
                       s->stream, s->misc.as<double>(), s->lduh.as<double>(), dx[0], dx[1], dx[2],
                       pde_has_source_or_ncp(s->cfg.pde) ? 1 : 0, s->cfg.pde == EXA_PDE_ADVECTION_NCP ? 0 : 1);
  });
  EXA_CHECK(hipGetLastError());
  std::vector<double> h(B3 * V);
  EXA_CHECK(hipMemcpyAsync(h.data(), s->lduh.p, sizeof(double) * B3 * V, hipMemcpyDeviceToHost, s->stream));
  EXA_TRY(sync_stream(s));
  soa_to_ref_cell(h.data(), lduh, B3, V);
  return EXA_OK;
}

template <int MODE>
static int run_update_cell(exa_solver* s, double* luh_ref, double* lduh_ref, const double* lFhbnd_ref,
                           const double dx[3], double dt, double* dt_out) {
  const int N = s->cfg.order, B = N + 1, B2 = B * B, B3 = B2 * B, V = exa::V;
  const int nd = s->cfg.nvar + s->cfg.npar;
  EXA_TRY(s->luh.ensure(sizeof(double) * B3 * V));
  EXA_TRY(s->lduh.ensure(sizeof(double) * B3 * V));
  EXA_TRY(s->ff.ensure(sizeof(double) * 6 * B2 * V));
  EXA_TRY(s->dt.ensure(sizeof(double) * 4));
  std::vector<double> h(std::max(B3 * V, 6 * B2 * V));
  if (luh_ref) {
    ref_to_soa_cell(luh_ref, h.data(), B3, nd);
    EXA_CHECK(hipMemcpy(s->luh.p, h.data(), sizeof(double) * B3 * V, hipMemcpyHostToDevice));
  }
  if (lduh_ref) {
    ref_to_soa_cell(lduh_ref, h.data(), B3, V);
    EXA_CHECK(hipMemcpy(s->lduh.p, h.data(), sizeof(double) * B3 * V, hipMemcpyHostToDevice));
  }
  if (lFhbnd_ref) {
    ref_to_soa_faces(lFhbnd_ref, h.data(), 6, B2, V);
    EXA_CHECK(hipMemcpy(s->ff.p, h.data(), sizeof(double) * 6 * B2 * V, hipMemcpyHostToDevice));
  }
  double hdt[2] = {dt, 0.0};
  unsigned long long inf_bits;
  const double big = 1.0e300;
  std::memcpy(&inf_bits, &big, sizeof(big));
  std::memcpy(&hdt[1], &inf_bits, sizeof(big));
  EXA_CHECK(hipMemcpy(s->dt.p, hdt, sizeof(hdt), hipMemcpyHostToDevice));
  double* dtp = s->dt.as<double>();
  auto* dtn = reinterpret_cast<unsigned long long*>(dtp + 1);
  if ((MODE & exa::UPD_DT) && exa_pde_generic(s->cfg.pde)) {
    hipLaunchKernelGGL(s->cfg.linear ? exa_pde_linear_update(s->cfg.pde, N, MODE) : exa_pde_update(s->cfg.pde, N, MODE),
                       dim3(1), dim3(round_up64(B3)), 0, s->stream,
                       s->luh.as<double>(), s->lduh.as<double>(), s->ff.as<double>(), dtp, dtn, dx[0], dx[1], dx[2],
                       s->cfg.gamma, s->cfg.cfl, launch_extras(s->cfg));
  } else EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    hipLaunchKernelGGL((exa::k_update<NN, MODE>), dim3(1), dim3(O::NT), 0, s->stream, s->luh.as<double>(),
                       s->lduh.as<double>(), s->ff.as<double>(), dtp, dtn, dx[0], dx[1], dx[2], s->cfg.gamma,
                       s->cfg.cfl, exa::Ens{});
  });
  EXA_CHECK(hipGetLastError());
  EXA_TRY(sync_stream(s));
  if (MODE & exa::UPD_UPDATE) {
    EXA_CHECK(hipMemcpy(h.data(), s->luh.p, sizeof(double) * B3 * V, hipMemcpyDeviceToHost));
    soa_to_ref_cell(h.data(), luh_ref, B3, nd);
  }
  if (MODE & exa::UPD_WRITE_LDUH) {
    EXA_CHECK(hipMemcpy(h.data(), s->lduh.p, sizeof(double) * B3 * V, hipMemcpyDeviceToHost));
    soa_to_ref_cell(h.data(), lduh_ref, B3, V);
  }
  if (MODE & exa::UPD_DT) {
    EXA_CHECK(hipMemcpy(hdt, s->dt.p, sizeof(hdt), hipMemcpyDeviceToHost));
    *dt_out = hdt[1];
    EXA_TRY(exa_check_dt(hdt[1], "stableTimeStepSize"));
  }
  return EXA_OK;
}

extern "C" int exa_aderdg_surface_integral(exa_solver* s, double* lduh, const double* lFhbnd, const double dx[3]) {
  EXA_ARG(s && lduh && lFhbnd && dx, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  if (s->cfg.linear)
    return run_update_cell<exa::UPD_SURFACE | exa::UPD_WRITE_LDUH | exa::UPD_LINEAR>(s, nullptr, lduh, lFhbnd, dx,
                                                                                      0.0, nullptr);
  return run_update_cell<exa::UPD_SURFACE | exa::UPD_WRITE_LDUH>(s, nullptr, lduh, lFhbnd, dx, 0.0, nullptr);
}

extern "C" int exa_aderdg_solution_update(exa_solver* s, double* luh, const double* lduh, double dt) {
  EXA_ARG(s && luh && lduh, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  const double dx[3] = {1.0, 1.0, 1.0};
  return run_update_cell<exa::UPD_UPDATE>(s, luh, const_cast<double*>(lduh), nullptr, dx, dt, nullptr);
}

extern "C" int exa_aderdg_stable_time_step_size(exa_solver* s, const double* luh, const double dx[3], double* dt) {
  EXA_ARG(s && luh && dx && dt, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  return run_update_cell<exa::UPD_DT>(s, const_cast<double*>(luh), nullptr, nullptr, dx, 0.0, dt);
}

extern "C" int exa_aderdg_riemann_solver(exa_solver* s, double* FL, double* FR, const double* QL, const double* QR,
                                         double dt, int normalNonZero) {
  (void)dt;
  EXA_ARG(s && FL && FR && QL && QR, "null pointer");
  EXA_ARG(normalNonZero >= 0 && normalNonZero < 3, "normalNonZero must be 0, 1 or 2");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  const int N = s->cfg.order, B = N + 1, B2 = B * B, V = exa::V, nd = s->cfg.nvar + s->cfg.npar;
  const int d = normalNonZero;
  // two cells: the left cell's face 2d+1 meets the right cell's face 2d
  std::vector<double> hq(2 * 6 * V * B2, 0.0), hf(2 * 6 * V * B2, 0.0);
  ref_to_soa_cell(QL, hq.data() + (0 * 6 + 2 * d + 1) * V * B2, B2, nd);
  ref_to_soa_cell(QR, hq.data() + (1 * 6 + 2 * d) * V * B2, B2, nd);
  ref_to_soa_cell(FL, hf.data() + (0 * 6 + 2 * d + 1) * V * B2, B2, V);
  ref_to_soa_cell(FR, hf.data() + (1 * 6 + 2 * d) * V * B2, B2, V);
  EXA_TRY(s->qf.ensure(sizeof(double) * hq.size()));
  EXA_TRY(s->ff.ensure(sizeof(double) * hf.size()));
  EXA_TRY(s->misc.ensure(sizeof(int) * 6));
  const int nbr[6] = {1, 0, 1, 0, 1, 0};
  EXA_CHECK(hipMemcpy(s->qf.p, hq.data(), sizeof(double) * hq.size(), hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(s->ff.p, hf.data(), sizeof(double) * hf.size(), hipMemcpyHostToDevice));
  EXA_CHECK(hipMemcpy(s->misc.p, nbr, sizeof(nbr), hipMemcpyHostToDevice));
  if (exa_pde_generic(s->cfg.pde)) {
    hipLaunchKernelGGL(s->cfg.linear ? exa_pde_linear_riemann(s->cfg.pde, N) : exa_pde_riemann(s->cfg.pde, N),
                       dim3(1), dim3(64), 0, s->stream, s->qf.as<double>(),
                       s->ff.as<double>(), s->misc.as<int>(), 2, s->cfg.gamma, (long long)(2 * d), 1LL,
                       launch_extras(s->cfg));
  } else EXA_ORDER_SWITCH(N, {
    hipLaunchKernelGGL((exa::k_riemann<NN>), dim3(1), dim3(64), 0, s->stream, s->qf.as<double>(), s->ff.as<double>(),
                       s->misc.as<int>(), 2, s->cfg.gamma, (long long)(2 * d), 1LL, exa::Ens{});
  });
  EXA_CHECK(hipGetLastError());
  EXA_TRY(sync_stream(s));
  EXA_CHECK(hipMemcpy(hf.data(), s->ff.p, sizeof(double) * hf.size(), hipMemcpyDeviceToHost));
  soa_to_ref_cell(hf.data() + (0 * 6 + 2 * d + 1) * V * B2, FL, B2, V);
  soa_to_ref_cell(hf.data() + (1 * 6 + 2 * d) * V * B2, FR, B2, V);
  return EXA_OK;
}

static bool bc_needs_euler(const exa_bc* bc) { return bc->kind == EXA_BC_REFLECTING || bc->kind == EXA_BC_EXACT; }
// Kernels of a registered PDE (exahype_amd_pde.h), by their launch signature.
using UserIcFn = void (*)(double*, int, int, double, double, double, double, double, double);
using UserBoundaryFn = void (*)(double*, double*, const int*, int, double*, double*, const double*, const double*,
                                double, double, double, double, double, double, int, int, double, exa::Ens);
using UserBoundaryRefFn = void (*)(double*, double*, const double*, const double*, double, double, double, double,
                                   double, double, double, double, int);

static bool pde_is_euler(int pde) { return pde == EXA_PDE_EULER || pde == EXA_PDE_EULER_GRAVITY; }

extern "C" int exa_aderdg_boundary_conditions(exa_solver* s, const exa_bc* bc, double* FOut, double* QOut,
                                              const double* FIn, const double* QIn, const double centre[3],
                                              const double size[3], double t, double dt, int faceIndex,
                                              int normalNonZero) {
  EXA_ARG(s && bc && FOut && QOut && FIn && QIn && centre && size, "null pointer");
  EXA_ARG(normalNonZero >= 0 && normalNonZero < 3, "normalNonZero must be 0, 1 or 2");
  EXA_ARG(faceIndex == 2 * normalNonZero || faceIndex == 2 * normalNonZero + 1, "faceIndex must be 2 normalNonZero + 0 or 1");
  const bool user = bc->kind == EXA_BC_USER;
  EXA_ARG((bc->kind >= EXA_BC_OUTFLOW && bc->kind <= EXA_BC_EXACT) || user, "boundary condition kind");
  EXA_ARG(!bc_needs_euler(bc) || pde_is_euler(s->cfg.pde), "reflecting and exact boundaries need an Euler PDE");
  EXA_ARG(!user || exa_pde_user(s->cfg.pde), "EXA_BC_USER needs a registered PDE");
  EXA_ARG(s->cfg.nvar == exa::V, "boundary values of 5 variables");
  EXA_ARG(!user || s->cfg.npar == 0, "a registered PDE carries no parameters");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  const int N = s->cfg.order, B2 = (N + 1) * (N + 1), V = exa::V, nd = s->cfg.nvar + s->cfg.npar;
  const size_t nq = (size_t)B2 * nd, nf = (size_t)B2 * V;
  EXA_TRY(s->qf.ensure(sizeof(double) * 2 * nq));
  EXA_TRY(s->ff.ensure(sizeof(double) * 2 * nf));
  double* dq = s->qf.as<double>();
  double* df = s->ff.as<double>();
  EXA_CHECK(hipMemcpyAsync(dq, QIn, sizeof(double) * nq, hipMemcpyHostToDevice, s->stream));
  EXA_CHECK(hipMemcpyAsync(df, FIn, sizeof(double) * nf, hipMemcpyHostToDevice, s->stream));
  if (user) {
    const auto k = reinterpret_cast<UserBoundaryRefFn>(
        const_cast<void*>(exa_pde_user_kernel(s->cfg.pde, N, EXA_USER_BOUNDARY_REF)));
    EXA_ARG(k, "the registered PDE has no kernels for this order");
    hipLaunchKernelGGL(k, dim3(1), dim3(64), 0, s->stream, df + nf, dq + nq, df, dq, centre[0], centre[1], centre[2],
                       size[0], size[1], size[2], t, dt, faceIndex);
    EXA_CHECK(hipGetLastError());
    EXA_CHECK(hipMemcpyAsync(QOut, dq + nq, sizeof(double) * nq, hipMemcpyDeviceToHost, s->stream));
    EXA_CHECK(hipMemcpyAsync(FOut, df + nf, sizeof(double) * nf, hipMemcpyDeviceToHost, s->stream));
    return sync_stream(s);
  }
  EXA_ORDER_SWITCH(N, {
    hipLaunchKernelGGL((exa::k_boundary_conditions_ref<NN>), dim3(1), dim3(64), 0, s->stream, df + nf, dq + nq, df,
                       dq, *bc, s->cfg.gamma, centre[0], centre[1], centre[2], size[0], size[1], size[2], t, dt,
                       faceIndex, normalNonZero, nd);
  });
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipMemcpyAsync(QOut, dq + nq, sizeof(double) * nq, hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipMemcpyAsync(FOut, df + nf, sizeof(double) * nf, hipMemcpyDeviceToHost, s->stream));
  return sync_stream(s);
}

extern "C" int exa_aderdg_adjust_solution(exa_solver* s, double* luh, const double centre[3], const double dx[3],
                                          double t, const exa_ic* ic) {
  EXA_ARG(s && luh && centre && dx && ic, "null pointer");
  EXA_ARG(t == 0.0, "only t=0 initial conditions are evaluated on the device");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  const int N = s->cfg.order, B = N + 1, B3 = B * B * B, V = exa::V, nd = s->cfg.nvar + s->cfg.npar;
  EXA_TRY(s->luh.ensure(sizeof(double) * B3 * V));
  EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    hipLaunchKernelGGL((exa::k_initial_conditions<NN>), dim3(1), dim3(O::NT), 0, s->stream, s->luh.as<double>(), *ic,
                       s->cfg.gamma, 1, 1, centre[0] - 0.5 * dx[0], centre[1] - 0.5 * dx[1],
                       centre[2] - 0.5 * dx[2], dx[0], dx[1], dx[2], exa::Ens{});
  });
  EXA_CHECK(hipGetLastError());
  EXA_TRY(sync_stream(s));
  std::vector<double> h(B3 * V);
  EXA_CHECK(hipMemcpy(h.data(), s->luh.p, sizeof(double) * B3 * V, hipMemcpyDeviceToHost));
  soa_to_ref_cell(h.data(), luh, B3, nd);
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// Deferred per-cell entry points.  The reference calls its kernels cell by
// cell (or face by face) inside one mesh traversal and consumes the results
// only in the next traversal: spaceTimePredictor + volumeIntegral in the
// prediction sweep (performPredictionAndVolumeIntegral,
// E/solvers/ADERDGSolver.cpp:1818-1946), riemannSolver in the merging sweep
// (solveRiemannProblemAtInterface, :2744-2825), surfaceIntegral +
// solutionUpdate in the update sweep (updateSolution, :2100-2179).  The
// batch_* calls copy their inputs and queue the work; exa_aderdg_batch_flush
// (at the traversal's end) runs every queued item of a kind in one launch and
// writes the outputs to the callers' arrays.

static int predictor_waves_per_simd(const exa_aderdg_config& c, int N, int dev);

static exa::InstParams cell_params(const exa_solver* s, const double dx[3], double dt) {
  exa::InstParams P{};
  P.dx[0] = dx[0];
  P.dx[1] = dx[1];
  P.dx[2] = dx[2];
  P.gamma = s->cfg.gamma;
  P.dt_cur = dt;
  const double big = 1.0e300;
  std::memcpy(&P.dt_next, &big, sizeof(big));
  return P;
}

extern "C" int exa_aderdg_batch_prediction(exa_solver* s, double* lQhbnd, double* lFhbnd, double* lduh,
                                           const double* luh, const double dx[3], double dt) {
  EXA_ARG(s && lQhbnd && lFhbnd && lduh && luh && dx, "null pointer");
  EXA_ARG(dt > 0.0, "dt must be positive");
  EXA_ARG(!exa_pde_generic(s->cfg.pde), "the batched entry points run the Euler PDEs");
  std::lock_guard<std::mutex> lk(s->mu);
  const int B = s->cfg.order + 1, B3 = B * B * B, nd = s->cfg.nvar + s->cfg.npar;
  const size_t off = s->bp_luh.size();
  s->bp_luh.resize(off + (size_t)exa::V * B3);
  ref_to_soa_cell(luh, s->bp_luh.data() + off, B3, nd);
  if (s->cfg.npar > 0) s->bp_par.insert(s->bp_par.end(), luh, luh + (size_t)B3 * nd);
  s->bp_inst.push_back(cell_params(s, dx, dt));
  s->bp.push_back({lQhbnd, lFhbnd, lduh});
  return EXA_OK;
}

extern "C" int exa_aderdg_batch_riemann(exa_solver* s, double* FL, double* FR, const double* QL, const double* QR,
                                        double dt, int normalNonZero) {
  EXA_ARG(s && FL && FR && QL && QR, "null pointer");
  EXA_ARG(normalNonZero >= 0 && normalNonZero < 3, "normalNonZero must be 0, 1 or 2");
  EXA_ARG(!exa_pde_generic(s->cfg.pde), "the batched entry points run the Euler PDEs");
  std::lock_guard<std::mutex> lk(s->mu);
  const int B = s->cfg.order + 1, B2 = B * B, V = exa::V, nd = s->cfg.nvar + s->cfg.npar;
  const size_t off = s->br_q.size();
  s->br_q.resize(off + 2 * (size_t)V * B2);
  s->br_f.resize(off + 2 * (size_t)V * B2);
  ref_to_soa_cell(QL, s->br_q.data() + off, B2, nd);
  ref_to_soa_cell(QR, s->br_q.data() + off + (size_t)V * B2, B2, nd);
  ref_to_soa_cell(FL, s->br_f.data() + off, B2, V);
  ref_to_soa_cell(FR, s->br_f.data() + off + (size_t)V * B2, B2, V);
  s->br_dir.push_back(normalNonZero);
  s->br_dt.push_back(dt);
  s->br.push_back({FL, FR});
  return EXA_OK;
}

extern "C" int exa_aderdg_batch_update(exa_solver* s, double* luh, double* lduh, const double* lFhbnd,
                                       const double dx[3], double dt) {
  EXA_ARG(s && luh && lduh && lFhbnd && dx, "null pointer");
  EXA_ARG(!exa_pde_generic(s->cfg.pde), "the batched entry points run the Euler PDEs");
  std::lock_guard<std::mutex> lk(s->mu);
  const int B = s->cfg.order + 1, B2 = B * B, B3 = B2 * B, V = exa::V, nd = s->cfg.nvar + s->cfg.npar;
  const size_t off = s->bu_luh.size(), offf = s->bu_ff.size();
  s->bu_luh.resize(off + (size_t)V * B3);
  s->bu_lduh.resize(off + (size_t)V * B3);
  s->bu_ff.resize(offf + (size_t)6 * V * B2);
  ref_to_soa_cell(luh, s->bu_luh.data() + off, B3, nd);
  ref_to_soa_cell(lduh, s->bu_lduh.data() + off, B3, V);
  ref_to_soa_faces(lFhbnd, s->bu_ff.data() + offf, 6, B2, V);
  s->bu_inst.push_back(cell_params(s, dx, dt));
  s->bu.push_back({luh, lduh});
  return EXA_OK;
}

extern "C" int exa_aderdg_batch_pending(exa_solver* s, long long counts[3]) {
  EXA_ARG(s && counts, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  counts[0] = (long long)s->bp.size();
  counts[1] = (long long)s->br.size();
  counts[2] = (long long)s->bu.size();
  return EXA_OK;
}

static void batch_clear(exa_solver* s) {
  s->bp.clear();
  s->bp_luh.clear();
  s->bp_par.clear();
  s->bp_inst.clear();
  s->br.clear();
  s->br_q.clear();
  s->br_f.clear();
  s->br_dir.clear();
  s->br_dt.clear();
  s->bu.clear();
  s->bu_luh.clear();
  s->bu_lduh.clear();
  s->bu_ff.clear();
  s->bu_inst.clear();
}

static int batch_run_predictions(exa_solver* s) {
  const int N = s->cfg.order, B = N + 1, B2 = B * B, B3 = B2 * B, V = exa::V, nd = s->cfg.nvar + s->cfg.npar;
  const long long n = (long long)s->bp.size();
  if (n == 0) return EXA_OK;
  std::vector<int> ci(n);
  for (long long c = 0; c < n; ++c) ci[c] = (int)c;
  EXA_TRY(s->b_luh.ensure(sizeof(double) * s->bp_luh.size()));
  EXA_TRY(s->b_lduh.ensure(sizeof(double) * n * V * B3));
  EXA_TRY(s->b_qf.ensure(sizeof(double) * n * 6 * V * B2));
  EXA_TRY(s->b_ff.ensure(sizeof(double) * n * 6 * V * B2));
  EXA_TRY(s->b_inst.ensure(sizeof(exa::InstParams) * n));
  EXA_TRY(s->b_cinst.ensure(sizeof(int) * n));
  EXA_TRY(s->b_iters.ensure(sizeof(int) * n));
  EXA_CHECK(hipMemcpyAsync(s->b_luh.p, s->bp_luh.data(), sizeof(double) * s->bp_luh.size(), hipMemcpyHostToDevice,
                           s->stream));
  EXA_CHECK(hipMemcpyAsync(s->b_inst.p, s->bp_inst.data(), sizeof(exa::InstParams) * n, hipMemcpyHostToDevice,
                           s->stream));
  EXA_CHECK(hipMemcpyAsync(s->b_cinst.p, ci.data(), sizeof(int) * n, hipMemcpyHostToDevice, s->stream));
  exa::Ens ens = launch_extras(s->cfg);
  ens.cell_inst = s->b_cinst.as<int>();
  ens.inst = s->b_inst.as<exa::InstParams>();
  exa::PredictorDebug nodbg;
  nodbg.lQi = nodbg.lFi = nodbg.lQhi = nodbg.lFhi = nullptr;
  EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    if (predictor_staged(s->cfg, s->device)) {
      int slots = 0;
      EXA_TRY(staged_slots(NN, s->device, n, &slots));
      EXA_TRY(s->b_stage.ensure(staged_slot_bytes(NN) * (size_t)slots));
      hipLaunchKernelGGL(exa_pred_staged_kernel(NN, false, s->cfg.pde == EXA_PDE_EULER_GRAVITY,
                                                s->cfg.precision == EXA_PREC_FP32),
                         dim3((unsigned)slots), dim3(O::NT),
                         (size_t)exa_pred_staged_lds_bytes(NN), s->stream, s->b_luh.as<double>(),
                         s->b_lduh.as<double>(), s->b_qf.as<double>(), s->b_ff.as<double>(), s->b_iters.as<int>(),
                         nullptr, 0.0, 0.0, 0.0, s->cfg.gamma, n, s->b_stage.as<double>(), nullptr, nodbg, ens,
                         s->cfg.fp64_sweeps);
    } else {
      const int wpe = predictor_waves_per_simd(s->cfg, NN, s->device);
      const int form = predictor_form(s->cfg, NN, s->device);
      PredFn kern = s->cfg.pde == EXA_PDE_EULER_GRAVITY ? predictor_kernel_src<NN>(false)
                                                          : predictor_kernel<NN>(false, s->cfg.precision, wpe, form);
      hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(O::NT), O::LDS_DOUBLES * sizeof(double), s->stream,
                         s->b_luh.as<double>(), s->b_lduh.as<double>(), s->b_qf.as<double>(), s->b_ff.as<double>(),
                         s->b_iters.as<int>(), nullptr, 0.0, 0.0, 0.0, s->cfg.gamma, s->cfg.fp64_sweeps, nullptr,
                         nodbg, ens);
    }
  });
  EXA_CHECK(hipGetLastError());
  std::vector<double> hq((size_t)n * 6 * V * B2), hf((size_t)n * 6 * V * B2), hd((size_t)n * V * B3);
  EXA_CHECK(hipMemcpyAsync(hq.data(), s->b_qf.p, sizeof(double) * hq.size(), hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipMemcpyAsync(hf.data(), s->b_ff.p, sizeof(double) * hf.size(), hipMemcpyDeviceToHost, s->stream));
  EXA_CHECK(hipMemcpyAsync(hd.data(), s->b_lduh.p, sizeof(double) * hd.size(), hipMemcpyDeviceToHost, s->stream));
  EXA_TRY(sync_stream(s));
  std::vector<double> parh;
  for (long long c = 0; c < n; ++c) {
    const auto& it = s->bp[c];
    soa_to_ref_faces(hq.data() + (size_t)c * 6 * V * B2, it.lQhbnd, 6, B2, nd);
    soa_to_ref_faces(hf.data() + (size_t)c * 6 * V * B2, it.lFhbnd, 6, B2, V);
    soa_to_ref_cell(hd.data() + (size_t)c * V * B3, it.lduh, B3, V);
    if (s->cfg.npar > 0) param_predictor(N, nd, s->bp_par.data() + (size_t)c * B3 * nd, it.lQhbnd, parh);
  }
  return EXA_OK;
}

static int batch_run_riemann(exa_solver* s) {
  const int N = s->cfg.order, B = N + 1, B2 = B * B, V = exa::V;
  const long long m = (long long)s->br.size();
  if (m == 0) return EXA_OK;
  EXA_TRY(s->b_qf.ensure(sizeof(double) * s->br_q.size()));
  EXA_TRY(s->b_ff.ensure(sizeof(double) * s->br_f.size()));
  EXA_TRY(s->b_dirs.ensure(sizeof(int) * m));
  EXA_CHECK(hipMemcpyAsync(s->b_qf.p, s->br_q.data(), sizeof(double) * s->br_q.size(), hipMemcpyHostToDevice,
                           s->stream));
  EXA_CHECK(hipMemcpyAsync(s->b_ff.p, s->br_f.data(), sizeof(double) * s->br_f.size(), hipMemcpyHostToDevice,
                           s->stream));
  EXA_CHECK(hipMemcpyAsync(s->b_dirs.p, s->br_dir.data(), sizeof(int) * m, hipMemcpyHostToDevice, s->stream));
  const unsigned blocks = (unsigned)((m * 64 + 255) / 256);
  EXA_ORDER_SWITCH(N, {
    hipLaunchKernelGGL((exa::k_riemann_list<NN>), dim3(blocks), dim3(256), 0, s->stream, s->b_qf.as<double>(),
                       s->b_ff.as<double>(), s->b_dirs.as<int>(), m, s->cfg.gamma);
  });
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipMemcpyAsync(s->br_f.data(), s->b_ff.p, sizeof(double) * s->br_f.size(), hipMemcpyDeviceToHost,
                           s->stream));
  EXA_TRY(sync_stream(s));
  for (long long i = 0; i < m; ++i) {
    soa_to_ref_cell(s->br_f.data() + (size_t)(2 * i) * V * B2, s->br[i].FL, B2, V);
    soa_to_ref_cell(s->br_f.data() + (size_t)(2 * i + 1) * V * B2, s->br[i].FR, B2, V);
  }
  return EXA_OK;
}

static int batch_run_updates(exa_solver* s) {
  const int N = s->cfg.order, B = N + 1, B2 = B * B, B3 = B2 * B, V = exa::V, nd = s->cfg.nvar + s->cfg.npar;
  const long long n = (long long)s->bu.size();
  if (n == 0) return EXA_OK;
  std::vector<int> ci(n);
  for (long long c = 0; c < n; ++c) ci[c] = (int)c;
  EXA_TRY(s->b_luh.ensure(sizeof(double) * s->bu_luh.size()));
  EXA_TRY(s->b_lduh.ensure(sizeof(double) * s->bu_lduh.size()));
  EXA_TRY(s->b_ff.ensure(sizeof(double) * s->bu_ff.size()));
  EXA_TRY(s->b_inst.ensure(sizeof(exa::InstParams) * n));
  EXA_TRY(s->b_cinst.ensure(sizeof(int) * n));
  EXA_CHECK(hipMemcpyAsync(s->b_luh.p, s->bu_luh.data(), sizeof(double) * s->bu_luh.size(), hipMemcpyHostToDevice,
                           s->stream));
  EXA_CHECK(hipMemcpyAsync(s->b_lduh.p, s->bu_lduh.data(), sizeof(double) * s->bu_lduh.size(),
                           hipMemcpyHostToDevice, s->stream));
  EXA_CHECK(hipMemcpyAsync(s->b_ff.p, s->bu_ff.data(), sizeof(double) * s->bu_ff.size(), hipMemcpyHostToDevice,
                           s->stream));
  EXA_CHECK(hipMemcpyAsync(s->b_inst.p, s->bu_inst.data(), sizeof(exa::InstParams) * n, hipMemcpyHostToDevice,
                           s->stream));
  EXA_CHECK(hipMemcpyAsync(s->b_cinst.p, ci.data(), sizeof(int) * n, hipMemcpyHostToDevice, s->stream));
  exa::Ens ens{s->b_cinst.as<int>(), s->b_inst.as<exa::InstParams>(), nullptr};
  EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    hipLaunchKernelGGL((exa::k_update<NN, exa::UPD_SURFACE | exa::UPD_UPDATE | exa::UPD_WRITE_LDUH>), dim3((unsigned)n),
                       dim3(O::NT), 0, s->stream, s->b_luh.as<double>(), s->b_lduh.as<double>(), s->b_ff.as<double>(),
                       nullptr, nullptr, 0.0, 0.0, 0.0, s->cfg.gamma, s->cfg.cfl, ens);
  });
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipMemcpyAsync(s->bu_luh.data(), s->b_luh.p, sizeof(double) * s->bu_luh.size(), hipMemcpyDeviceToHost,
                           s->stream));
  EXA_CHECK(hipMemcpyAsync(s->bu_lduh.data(), s->b_lduh.p, sizeof(double) * s->bu_lduh.size(),
                           hipMemcpyDeviceToHost, s->stream));
  EXA_TRY(sync_stream(s));
  for (long long c = 0; c < n; ++c) {
    soa_to_ref_cell(s->bu_luh.data() + (size_t)c * V * B3, s->bu[c].luh, B3, nd);
    soa_to_ref_cell(s->bu_lduh.data() + (size_t)c * V * B3, s->bu[c].lduh, B3, V);
  }
  return EXA_OK;
}

extern "C" int exa_aderdg_batch_flush(exa_solver* s) {
  EXA_ARG(s, "null pointer");
  std::lock_guard<std::mutex> lk(s->mu);
  EXA_CHECK(hipSetDevice(s->device));
  int r = batch_run_predictions(s);
  if (r == EXA_OK) r = batch_run_riemann(s);
  if (r == EXA_OK) r = batch_run_updates(s);
  batch_clear(s);
  return r;
}

// ---------------------------------------------------------------------------
// device-resident grid
struct exa_grid {
  exa_solver* s = nullptr;
  int nx = 0, ny = 0, nz = 0;
  long long nc = 0;
  double off[3] = {0, 0, 0}, width[3] = {1, 1, 1}, dx[3] = {1, 1, 1};
  DevBuf luh, lduh, qf, ff, iters, nbr, state;
  DevBuf stage;    // staged predictor: one staging slot per resident workgroup
  // domain boundaries (exa_grid_set_boundary): the faces of the non-periodic
  // axes, cell * 6 + face, and their outside states and fluxes
  int periodic[3] = {1, 1, 1};
  exa_bc bc{};
  int nb = 0;
  DevBuf bfaces, bq, bf;
  DevBuf sample;   // Cartesian-vertex samples (output)
  int nslots = 0;
  bool staged = false;  // decided at creation (the slots are allocated then)
  hipStream_t stream = nullptr;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool timing = false;
  bool dt_valid = false;  // a dt has been computed or set (sync checks it from then on)
  bool fused = false;     // fused time stepping (exa_grid_set_fused_time_stepping)
  double fuse_w = 1.0;    // its fuse-algorithmic-steps-factor
  long long version = 0;  // bumped whenever luh changes (steps, uploads, initial conditions)
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[3];
  double times[3] = {0, 0, 0};
  long long counts[3] = {0, 0, 0};
  double* dt_cur() const { return state.as<double>(); }
  unsigned long long* dt_next() const { return reinterpret_cast<unsigned long long*>(state.as<double>() + 1); }
  double* time() const { return state.as<double>() + 2; }
  unsigned long long* iter_sum() const { return reinterpret_cast<unsigned long long*>(state.as<double>() + 3); }
  double* dt_pred() const { return state.as<double>() + 4; }  // fused stepping: the optimistic dt
  unsigned long long* reruns() const { return reinterpret_cast<unsigned long long*>(state.as<double>() + 5); }
};

// +d neighbours (nbr[d * nc + c]) and -d neighbours (nbr[(3 + d) * nc + c]),
// wrapped on periodic axes and -1 past the first / last cell of the others,
// and the list of domain faces of the non-periodic axes (cell * 6 + face).
static hipError_t grid_neighbours(exa_grid* g) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int n[3] = {nx, ny, nz};
  std::vector<int> nbr(6 * g->nc), bfaces;
  for (long long c = 0; c < g->nc; ++c) {
    const int ix = (int)(c % nx), iy = (int)((c / nx) % ny), iz = (int)(c / ((long long)nx * ny));
    const int id[3] = {ix, iy, iz};
    for (int d = 0; d < 3; ++d) {
      int j[3] = {ix, iy, iz};
      j[d] = (id[d] + 1) % n[d];
      const bool edge = id[d] == n[d] - 1 && !g->periodic[d];
      nbr[d * g->nc + c] = edge ? -1 : (int)(((long long)j[2] * ny + j[1]) * nx + j[0]);
      j[d] = (id[d] + n[d] - 1) % n[d];
      const bool first = id[d] == 0 && !g->periodic[d];
      nbr[(3 + d) * g->nc + c] = first ? -1 : (int)(((long long)j[2] * ny + j[1]) * nx + j[0]);
      if (!g->periodic[d]) {
        if (id[d] == 0) bfaces.push_back((int)(c * 6 + 2 * d));
        if (id[d] == n[d] - 1) bfaces.push_back((int)(c * 6 + 2 * d + 1));
      }
    }
  }
  hipError_t e = hipMemcpy(g->nbr.p, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice);
  g->nb = (int)bfaces.size();
  if (e == hipSuccess && g->nb > 0) {
    const int B2 = (g->s->cfg.order + 1) * (g->s->cfg.order + 1);
    if (g->bfaces.ensure(sizeof(int) * g->nb) != EXA_OK || g->bq.ensure(sizeof(double) * g->nb * exa::V * B2) != EXA_OK ||
        g->bf.ensure(sizeof(double) * g->nb * exa::V * B2) != EXA_OK)
      return hipErrorOutOfMemory;
    e = hipMemcpy(g->bfaces.p, bfaces.data(), sizeof(int) * g->nb, hipMemcpyHostToDevice);
  }
  return e;
}

extern "C" int exa_grid_create(exa_solver* s, int nx, int ny, int nz, const double offset[3], const double width[3],
                               exa_grid** out) {
  EXA_ARG(s && out && nx > 0 && ny > 0 && nz > 0, "grid dimensions");
  EXA_ARG(s->cfg.npar == 0, "the device grid carries no material parameters (npar must be 0)");
  EXA_CHECK(hipSetDevice(s->device));
  exa_grid* g = new exa_grid();
  g->s = s;
  g->nx = nx;
  g->ny = ny;
  g->nz = nz;
  g->nc = (long long)nx * ny * nz;
  for (int d = 0; d < 3; ++d) {
    g->off[d] = offset ? offset[d] : 0.0;
    g->width[d] = width ? width[d] : 1.0;
  }
  g->dx[0] = g->width[0] / nx;
  g->dx[1] = g->width[1] / ny;
  g->dx[2] = g->width[2] / nz;
  const int B = s->cfg.order + 1, B2 = B * B, B3 = B2 * B, V = exa::V;
  int r = EXA_OK;
  if ((r = g->luh.ensure(sizeof(double) * g->nc * V * B3)) != EXA_OK ||
      (r = g->lduh.ensure(sizeof(double) * g->nc * V * B3)) != EXA_OK ||
      (r = g->qf.ensure(sizeof(double) * g->nc * 6 * V * B2)) != EXA_OK ||
      (r = g->ff.ensure(sizeof(double) * g->nc * 6 * V * B2)) != EXA_OK ||
      (r = g->iters.ensure(sizeof(int) * g->nc)) != EXA_OK || (r = g->nbr.ensure(sizeof(int) * 6 * g->nc)) != EXA_OK ||
      (r = g->state.ensure(sizeof(double) * 8)) != EXA_OK) {
    delete g;
    return r;
  }
  g->staged = predictor_staged(s->cfg, s->device);
  if (g->staged) {
    if ((r = staged_slots(s->cfg.order, s->device, g->nc, &g->nslots)) != EXA_OK ||
        (r = g->stage.ensure(staged_slot_bytes(s->cfg.order) * (size_t)g->nslots)) != EXA_OK) {
      delete g;
      return r;
    }
  }
  hipError_t e = grid_neighbours(g);
  double st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double big = 1.0e300;
  std::memcpy(&st[1], &big, sizeof(big));
  if (e == hipSuccess) e = hipMemcpy(g->state.p, st, sizeof(st), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(g->iters.p, 0, sizeof(int) * g->nc);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    exa_set_error("grid setup: %s", hipGetErrorString(e));
    delete g;
    return EXA_ERR_HIP;
  }
  *out = g;
  return EXA_OK;
}

static void grid_clear_events(exa_grid* g) {
  for (auto& v : g->ev)
    for (auto& p : v) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  for (auto& v : g->ev) v.clear();
}

int exa_grid_view(exa_grid* g, ExaGridView* v) {
  EXA_ARG(g && v, "null pointer");
  v->luh = g->luh.as<double>();
  v->nc = g->nc;
  v->nx = g->nx;
  v->ny = g->ny;
  v->nz = g->nz;
  v->order = g->s->cfg.order;
  v->device = g->s->device;
  v->stream = g->stream;
  v->version = g->version;
  v->periodic = g->periodic[0] && g->periodic[1] && g->periodic[2];
  return EXA_OK;
}

extern "C" int exa_grid_destroy(exa_grid* g) {
  if (!g) return EXA_OK;
  (void)hipSetDevice(g->s->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  exa_grid_limiter_release(g);
  grid_clear_events(g);
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return EXA_OK;
}

extern "C" int exa_grid_num_cells(const exa_grid* g, long long* n) {
  EXA_ARG(g && n, "null pointer");
  *n = g->nc;
  return EXA_OK;
}

static int grid_transfer(exa_grid* g, double* host, bool upload) {
  const int B = g->s->cfg.order + 1, B3 = B * B * B, V = exa::V;
  const size_t per = (size_t)V * B3;
  const long long chunk = 4096;
  std::vector<double> buf(per * chunk);
  for (long long c0 = 0; c0 < g->nc; c0 += chunk) {
    const long long n = std::min(chunk, g->nc - c0);
    double* dev = g->luh.as<double>() + c0 * per;
    if (upload) {
      for (long long c = 0; c < n; ++c) ref_to_soa_cell(host + (c0 + c) * per, buf.data() + c * per, B3, V);
      EXA_CHECK(hipMemcpyAsync(dev, buf.data(), sizeof(double) * per * n, hipMemcpyHostToDevice, g->stream));
      EXA_CHECK(hipStreamSynchronize(g->stream));
    } else {
      EXA_CHECK(hipMemcpyAsync(buf.data(), dev, sizeof(double) * per * n, hipMemcpyDeviceToHost, g->stream));
      EXA_CHECK(hipStreamSynchronize(g->stream));
      for (long long c = 0; c < n; ++c) soa_to_ref_cell(buf.data() + c * per, host + (c0 + c) * per, B3, V);
    }
  }
  return EXA_OK;
}

extern "C" int exa_grid_upload(exa_grid* g, const double* luh) {
  EXA_ARG(g && luh, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  ++g->version;
  return grid_transfer(g, const_cast<double*>(luh), true);
}

extern "C" int exa_grid_download(exa_grid* g, double* luh) {
  EXA_ARG(g && luh, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  return grid_transfer(g, luh, false);
}

extern "C" int exa_grid_sample_cartesian(exa_grid* g, double* out) {
  EXA_ARG(g && out, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  const int N = g->s->cfg.order, B = N + 1, B3 = B * B * B;
  const size_t n = (size_t)g->nc * B3 * exa::V;
  EXA_TRY(g->sample.ensure(sizeof(double) * n));
  EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    hipLaunchKernelGGL((exa::k_sample_cartesian<NN>), dim3((unsigned)g->nc), dim3(O::NT), 0, g->stream,
                       g->luh.as<double>(), g->sample.as<double>());
  });
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipMemcpyAsync(out, g->sample.p, sizeof(double) * n, hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_grid_set_initial_conditions(exa_grid* g, const exa_ic* ic) {
  const bool user = g && exa_pde_user(g->s->cfg.pde);
  EXA_ARG(g && (ic || user), "null pointer (only a registered PDE takes its own initial conditions)");
  EXA_CHECK(hipSetDevice(g->s->device));
  ++g->version;
  const int N = g->s->cfg.order;
  if (user && !ic) {
    const auto k = reinterpret_cast<UserIcFn>(const_cast<void*>(exa_pde_user_kernel(g->s->cfg.pde, N, EXA_USER_IC)));
    EXA_ARG(k, "the registered PDE has no kernels for this order");
    hipLaunchKernelGGL(k, dim3((unsigned)g->nc), dim3(round_up64((N + 1) * (N + 1) * (N + 1))), 0, g->stream,
                       g->luh.as<double>(), g->nx, g->ny, g->off[0], g->off[1], g->off[2], g->dx[0], g->dx[1],
                       g->dx[2]);
    EXA_CHECK(hipGetLastError());
    EXA_CHECK(hipStreamSynchronize(g->stream));
    return EXA_OK;
  }
  EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    hipLaunchKernelGGL((exa::k_initial_conditions<NN>), dim3((unsigned)g->nc), dim3(O::NT), 0, g->stream,
                       g->luh.as<double>(), *ic, g->s->cfg.gamma, g->nx, g->ny, g->off[0], g->off[1], g->off[2],
                       g->dx[0], g->dx[1], g->dx[2], exa::Ens{});
  });
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

template <int MODE>
static int launch_update(exa_grid* g) {
  const int N = g->s->cfg.order;
  if (exa_pde_generic(g->s->cfg.pde)) {
    PdeUpdateFn fn = g->s->cfg.linear
                         ? exa_pde_linear_update(g->s->cfg.pde, N, (MODE & exa::UPD_SURFACE) ? MODE | exa::UPD_LINEAR : MODE)
                         : exa_pde_update(g->s->cfg.pde, N, MODE);
    hipLaunchKernelGGL(fn, dim3((unsigned)g->nc), dim3(round_up64((N + 1) * (N + 1) * (N + 1))),
                       0, g->stream, g->luh.as<double>(), g->lduh.as<double>(), g->ff.as<double>(), g->dt_cur(),
                       g->dt_next(), g->dx[0], g->dx[1], g->dx[2], g->s->cfg.gamma, g->s->cfg.cfl,
                       launch_extras(g->s->cfg));
    EXA_CHECK(hipGetLastError());
    return EXA_OK;
  }
  EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    hipLaunchKernelGGL((exa::k_update<NN, MODE>), dim3((unsigned)g->nc), dim3(O::NT), 0, g->stream,
                       g->luh.as<double>(), g->lduh.as<double>(), g->ff.as<double>(), g->dt_cur(), g->dt_next(),
                       g->dx[0], g->dx[1], g->dx[2], g->s->cfg.gamma, g->s->cfg.cfl, exa::Ens{});
  });
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

extern "C" int exa_grid_compute_dt(exa_grid* g, double* dt) {
  EXA_ARG(g, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  hipLaunchKernelGGL(exa::k_dt_reset, dim3(1), dim3(1), 0, g->stream, g->dt_next());
  EXA_TRY(launch_update<exa::UPD_DT>(g));
  hipLaunchKernelGGL(exa::k_dt_take, dim3(1), dim3(1), 0, g->stream, g->dt_cur(), g->dt_next());
  hipLaunchKernelGGL(exa::k_dt_reset, dim3(1), dim3(1), 0, g->stream, g->dt_next());
  if (g->fused)  // fused stepping starts again from this state's stable dt
    hipLaunchKernelGGL(exa::k_dt_fused_start, dim3(1), dim3(1), 0, g->stream, g->dt_cur(), g->dt_pred(), g->fuse_w);
  EXA_CHECK(hipGetLastError());
  double h = 0.0;
  EXA_CHECK(hipMemcpyAsync(&h, g->dt_cur(), sizeof(double), hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  if (dt) *dt = h;
  g->dt_valid = true;
  return exa_check_dt(h, "exa_grid_compute_dt");
}

extern "C" int exa_grid_set_dt(exa_grid* g, double dt) {
  EXA_ARG(g && dt > 0.0, "dt must be positive");
  EXA_CHECK(hipSetDevice(g->s->device));
  EXA_CHECK(hipMemcpyAsync(g->dt_cur(), &dt, sizeof(double), hipMemcpyHostToDevice, g->stream));
  // fused stepping: the prescribed dt is also the optimistic one
  if (g->fused) EXA_CHECK(hipMemcpyAsync(g->dt_pred(), &dt, sizeof(double), hipMemcpyHostToDevice, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  g->dt_valid = true;
  return EXA_OK;
}

extern "C" int exa_grid_set_fused_time_stepping(exa_grid* g, int enable, double weight) {
  EXA_ARG(g, "null pointer");
  EXA_ARG(!enable || (weight > 0.0 && weight <= 1.0), "fuse-algorithmic-steps-factor must lie in (0, 1]");
  EXA_CHECK(hipSetDevice(g->s->device));
  g->fused = enable != 0;
  g->fuse_w = enable ? weight : 1.0;
  const double zero = 0.0;
  const unsigned long long nz = 0;
  EXA_CHECK(hipMemcpyAsync(g->dt_pred(), &zero, sizeof(zero), hipMemcpyHostToDevice, g->stream));
  EXA_CHECK(hipMemcpyAsync(g->reruns(), &nz, sizeof(nz), hipMemcpyHostToDevice, g->stream));
  // a dt in hand is taken as the current state's stable dt
  if (g->fused && g->dt_valid)
    hipLaunchKernelGGL(exa::k_dt_fused_start, dim3(1), dim3(1), 0, g->stream, g->dt_cur(), g->dt_pred(), g->fuse_w);
  EXA_CHECK(hipGetLastError());
  EXA_CHECK(hipStreamSynchronize(g->stream));
  // the captured step ends with the other time-step kernel
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->gexec = nullptr;
  g->graph = nullptr;
  return EXA_OK;
}

extern "C" int exa_grid_fused_reruns(exa_grid* g, long long* reruns) {
  EXA_ARG(g && reruns, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  unsigned long long h = 0;
  EXA_CHECK(hipMemcpyAsync(&h, g->reruns(), sizeof(h), hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  *reruns = (long long)h;
  return EXA_OK;
}

// Occupancy target of the predictor: measured per order and precision
// (two waves per SIMD trade a few spilled registers for latency hiding).
static int predictor_waves_per_simd(const exa_aderdg_config& c, int N, int dev) {
  if (c.waves_per_simd == 1 || c.waves_per_simd == 2) return c.waves_per_simd;
  const int tw = g_tuned[dev & 63][N][c.precision & 3].wpe;
  if (tw) return tw;
  // measured at 81^3 cells: two waves per SIMD is 1.6x faster at N=5
  return N <= 5 ? 2 : 1;
}

// The register-form predictor a solver's grids launch (Euler, orders 1..7,
// the configured precision, occupancy and contraction form), for other units.
int exa_solver_predictor(const exa_solver* s, PredFn* fn, int* threads, size_t* lds) {
  const int N = s->cfg.order;
  EXA_ARG(s->cfg.pde == EXA_PDE_EULER && !s->cfg.linear, "the Euler PDE (nonlinear kernels)");
  EXA_ARG(N >= 1 && N <= 7, "orders 1..7 (the register-form predictor)");
  EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    const int wpe = predictor_waves_per_simd(s->cfg, NN, s->device);
    const int form = predictor_form(s->cfg, NN, s->device);
    // the adaptive mesh is mostly uniform away from its features: FP32
    // sweeps skip the cells they cannot improve
    *fn = predictor_kernel<NN>(false, s->cfg.precision, wpe, form, true);
    *threads = O::NT;
    *lds = O::LDS_DOUBLES * sizeof(double);
  });
  return *fn ? EXA_OK : EXA_ERR_UNSUPPORTED;
}
const exa_aderdg_config* exa_solver_cfg(const exa_solver* s) { return &s->cfg; }
int exa_solver_device(const exa_solver* s) { return s->device; }

// Extras of a grid's launches: a point source is evaluated per node from the
// grid's geometry and the device-held time of the step.
static exa::Ens grid_extras(const exa_grid* g) {
  exa::Ens e = launch_extras(g->s->cfg);
  if (g->s->cfg.point_source) {
    e.ps.mode = 2;
    for (int d = 0; d < 3; ++d) e.ps.off[d] = g->off[d];
    e.ps.nx = g->nx;
    e.ps.ny = g->ny;
  }
  return e;
}

// Riemann problems, surface integral, update and CFL minimum of every cell in
// one launch (k_riemann_update), after a predictor launched with
// Ens::fold_update.  box: nx, ny, nz and the periodic axes of a uniform grid
// (neighbours computed in the kernel), or null for the table.
static int launch_riemann_update(int N, double* luh, const double* lduh, const double* qf, const double* ff,
                                 const int* nbr, long long nc, const int* box, const double* dt_cur,
                                 unsigned long long* dt_next, const double dx[3], double gamma, double cfl,
                                 hipStream_t st, const exa::Ens& en) {
  EXA_ARG(en.fold_update, "the fused update needs the folded predictor output");
  EXA_ARG(nc < (1LL << 32) - 1, "too many cells for one launch");
  const int nx = box ? box[0] : 0, ny = box ? box[1] : 0, nz = box ? box[2] : 0, per = box ? box[3] : 0;
  EXA_ORDER_SWITCH(N, {
    hipLaunchKernelGGL((exa::k_riemann_update<NN>), dim3((unsigned)nc), dim3(exa::RiemannUpdate<NN>::NT), 0, st,
                       luh, lduh, qf, ff, nbr, nc, nx, ny, nz, per, dt_cur, dt_next, dx[0], dx[1], dx[2], gamma, cfl,
                       en);
  });
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

static int enqueue_step(exa_grid* g, bool timed) {
  const int N = g->s->cfg.order;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto tstart = [&](int) -> int {
    if (!timed) return EXA_OK;
    EXA_CHECK(hipEventCreate(&e0));
    EXA_CHECK(hipEventCreate(&e1));
    EXA_CHECK(hipEventRecord(e0, g->stream));
    return EXA_OK;
  };
  auto tstop = [&](int phase) -> int {
    if (!timed) return EXA_OK;
    EXA_CHECK(hipEventRecord(e1, g->stream));
    g->ev[phase].emplace_back(e0, e1);
    return EXA_OK;
  };
  EXA_TRY(tstart(0));
  exa::PredictorDebug nodbg;
  nodbg.lQi = nodbg.lFi = nodbg.lQhi = nodbg.lFhi = nullptr;
  const bool generic = exa_pde_generic(g->s->cfg.pde);
  const bool linear = g->s->cfg.linear != 0;
  if (generic) {
    hipLaunchKernelGGL(linear ? exa_pde_linear_predictor(g->s->cfg.pde, N, false, g->s->cfg.point_source != 0)
                              : exa_pde_predictor(g->s->cfg.pde, N, false),
                       dim3((unsigned)g->nc), dim3(round_up64((N + 1) * (N + 1) * (N + 1))),
                       (size_t)(linear ? exa_pde_linear_lds_bytes(N) : exa_pde_predictor_lds_bytes(N)),
                       g->stream, g->luh.as<double>(), g->lduh.as<double>(), g->qf.as<double>(), g->ff.as<double>(),
                       g->iters.as<int>(), g->dt_cur(), g->dx[0], g->dx[1], g->dx[2], g->s->cfg.gamma,
                       g->iter_sum(), nodbg, grid_extras(g));
  } else EXA_ORDER_SWITCH(N, {
    using O = exa::Ord<NN>;
    exa::Ens ex = launch_extras(g->s->cfg);
    ex.fold_update = 1;  // k_riemann_update follows
    if (g->staged) {
      hipLaunchKernelGGL(exa_pred_staged_kernel(NN, false, g->s->cfg.pde == EXA_PDE_EULER_GRAVITY,
                                                g->s->cfg.precision == EXA_PREC_FP32),
                         dim3((unsigned)g->nslots), dim3(O::NT),
                         (size_t)exa_pred_staged_lds_bytes(NN), g->stream, g->luh.as<double>(), g->lduh.as<double>(),
                         g->qf.as<double>(), g->ff.as<double>(), g->iters.as<int>(), g->dt_cur(), g->dx[0],
                         g->dx[1], g->dx[2], g->s->cfg.gamma, g->nc, g->stage.as<double>(), g->iter_sum(), nodbg,
                         ex, g->s->cfg.fp64_sweeps);
    } else {
      const int wpe = predictor_waves_per_simd(g->s->cfg, NN, g->s->device);
      const int form = predictor_form(g->s->cfg, NN, g->s->device);
      PredFn kern = g->s->cfg.pde == EXA_PDE_EULER_GRAVITY ? predictor_kernel_src<NN>(false)
                                                             : predictor_kernel<NN>(false, g->s->cfg.precision, wpe, form);
      hipLaunchKernelGGL(kern, dim3((unsigned)g->nc), dim3(O::NT), O::LDS_DOUBLES * sizeof(double), g->stream,
                         g->luh.as<double>(), g->lduh.as<double>(), g->qf.as<double>(), g->ff.as<double>(),
                         g->iters.as<int>(), g->dt_cur(), g->dx[0], g->dx[1], g->dx[2], g->s->cfg.gamma,
                         g->s->cfg.fp64_sweeps, g->iter_sum(), nodbg, ex);
    }
  });
  EXA_CHECK(hipGetLastError());
  EXA_TRY(tstop(0));
  EXA_TRY(tstart(1));
  const long long waves = 3LL * g->nc;
  const unsigned blocks = (unsigned)((waves * 64 + 255) / 256);
  if (generic) {
    hipLaunchKernelGGL(linear ? exa_pde_linear_riemann(g->s->cfg.pde, N) : exa_pde_riemann(g->s->cfg.pde, N),
                       dim3(blocks), dim3(256), 0, g->stream, g->qf.as<double>(),
                       g->ff.as<double>(), g->nbr.as<int>(), (int)g->nc, g->s->cfg.gamma, 0LL, waves,
                       launch_extras(g->s->cfg));
    if (g->nb > 0) {  // only a registered PDE has closed axes here
      const auto k = reinterpret_cast<UserBoundaryFn>(
          const_cast<void*>(exa_pde_user_kernel(g->s->cfg.pde, N, EXA_USER_BOUNDARY_FACES)));
      EXA_ARG(k, "the PDE has no boundary kernel for this order");
      hipLaunchKernelGGL(k, dim3((unsigned)((g->nb * 64LL + 255) / 256)), dim3(256), 0, g->stream,
                         g->qf.as<double>(), g->ff.as<double>(), g->bfaces.as<int>(), g->nb, g->bq.as<double>(),
                         g->bf.as<double>(), g->dt_cur(), g->time(), g->off[0], g->off[1], g->off[2], g->dx[0],
                         g->dx[1], g->dx[2], g->nx, g->ny, g->s->cfg.gamma, launch_extras(g->s->cfg));
    }
  } else EXA_ORDER_SWITCH(N, {
    // domain faces first: the fused kernel takes their fluxes from ff
    if (g->nb > 0)
      hipLaunchKernelGGL((exa::k_boundary_faces<NN>), dim3((unsigned)((g->nb * 64LL + 255) / 256)), dim3(256), 0,
                         g->stream, g->qf.as<double>(), g->ff.as<double>(), g->bfaces.as<int>(), g->nb,
                         g->bq.as<double>(), g->bf.as<double>(), g->bc, g->s->cfg.gamma, g->dt_cur(), g->time(),
                         g->off[0], g->off[1], g->off[2], g->dx[0], g->dx[1], g->dx[2], g->nx, g->ny);
  });
  EXA_CHECK(hipGetLastError());
  EXA_TRY(tstop(1));
  EXA_TRY(tstart(2));
  if (generic) {
    EXA_TRY(launch_update<exa::UPD_SURFACE | exa::UPD_UPDATE | exa::UPD_DT>(g));
  } else {
    const int box[4] = {g->nx, g->ny, g->nz, (g->periodic[0] ? 1 : 0) | (g->periodic[1] ? 2 : 0) | (g->periodic[2] ? 4 : 0)};
    exa::Ens ex{};
    ex.fold_update = 1;
    EXA_TRY(launch_riemann_update(N, g->luh.as<double>(), g->lduh.as<double>(), g->qf.as<double>(),
                                  g->ff.as<double>(), g->nbr.as<int>(), g->nc, box, g->dt_cur(), g->dt_next(), g->dx,
                                  g->s->cfg.gamma, g->s->cfg.cfl, g->stream, ex));
  }
  EXA_TRY(tstop(2));
  if (g->fused)
    hipLaunchKernelGGL(exa::k_dt_advance_fused, dim3(1), dim3(1), 0, g->stream, g->dt_cur(), g->dt_next(), g->time(),
                       g->dt_pred(), g->reruns(), g->fuse_w);
  else
    hipLaunchKernelGGL(exa::k_dt_advance, dim3(1), dim3(1), 0, g->stream, g->dt_cur(), g->dt_next(), g->time());
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

extern "C" int exa_grid_set_boundary(exa_grid* g, const int periodic[3], const exa_bc* bc) {
  EXA_ARG(g && periodic, "null pointer");
  const bool all = periodic[0] && periodic[1] && periodic[2];
  if (!all) {
    EXA_ARG(bc, "a grid with a non-periodic axis needs boundary conditions");
    if (exa_pde_user(g->s->cfg.pde)) {
      EXA_ARG(bc->kind == EXA_BC_USER, "a registered PDE takes its own boundaryValues (EXA_BC_USER)");
    } else {
      EXA_ARG(bc->kind >= EXA_BC_OUTFLOW && bc->kind <= EXA_BC_EXACT, "boundary condition kind");
      EXA_ARG(pde_is_euler(g->s->cfg.pde) && !g->s->cfg.linear,
              "domain boundaries of the device grid are implemented for the Euler PDEs and registered PDEs");
    }
  }
  EXA_CHECK(hipSetDevice(g->s->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  for (int d = 0; d < 3; ++d) g->periodic[d] = periodic[d] ? 1 : 0;
  if (bc) g->bc = *bc;
  const hipError_t e = grid_neighbours(g);
  if (e != hipSuccess) {
    exa_set_error("grid boundary setup: %s", hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? EXA_ERR_NOMEM : EXA_ERR_HIP;
  }
  // the captured step holds the old neighbour list and boundary kernel
  if (g->gexec) (void)hipGraphExecDestroy(g->gexec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->gexec = nullptr;
  g->graph = nullptr;
  return EXA_OK;
}

extern "C" int exa_grid_use_graph(exa_grid* g, int enable) {
  EXA_ARG(g, "null pointer");
  g->use_graph = enable != 0;
  return EXA_OK;
}

extern "C" int exa_grid_enable_timing(exa_grid* g, int enable) {
  EXA_ARG(g, "null pointer");
  g->timing = enable != 0;
  return EXA_OK;
}

extern "C" int exa_grid_step(exa_grid* g, int nsteps) {
  EXA_ARG(g && nsteps >= 0, "nsteps");
  EXA_CHECK(hipSetDevice(g->s->device));
  g->version += nsteps;
  if (g->use_graph && !g->timing) {
    if (!g->gexec) {
      EXA_CHECK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
      int r = enqueue_step(g, false);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(g->stream, &graph);
      if (r != EXA_OK) return r;
      EXA_CHECK(e);
      g->graph = graph;
      EXA_CHECK(hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < nsteps; ++i) EXA_CHECK(hipGraphLaunch(g->gexec, g->stream));
    return EXA_OK;
  }
  for (int i = 0; i < nsteps; ++i) EXA_TRY(enqueue_step(g, g->timing));
  return EXA_OK;
}

extern "C" int exa_grid_sync(exa_grid* g, double* time, double* next_dt) {
  EXA_ARG(g, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  double st[3];
  EXA_CHECK(hipMemcpyAsync(st, g->state.p, sizeof(st), hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  EXA_CHECK(hipGetLastError());
  if (time) *time = st[2];
  if (next_dt) *next_dt = st[0];
  return g->dt_valid ? exa_check_dt(st[0], "exa_grid_sync") : EXA_OK;
}

extern "C" int exa_grid_picard_iterations(exa_grid* g, int* iters) {
  EXA_ARG(g && iters, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  EXA_CHECK(hipMemcpyAsync(iters, g->iters.p, sizeof(int) * g->nc, hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  return EXA_OK;
}

extern "C" int exa_grid_total_picard_iterations(exa_grid* g, long long* total) {
  EXA_ARG(g && total, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  unsigned long long h = 0, zero = 0;
  EXA_CHECK(hipMemcpyAsync(&h, g->iter_sum(), sizeof(h), hipMemcpyDeviceToHost, g->stream));
  EXA_CHECK(hipMemcpyAsync(g->iter_sum(), &zero, sizeof(zero), hipMemcpyHostToDevice, g->stream));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  *total = (long long)h;
  return EXA_OK;
}

extern "C" int exa_grid_kernel_times(exa_grid* g, double times[3], long long counts[3]) {
  EXA_ARG(g && times && counts, "null pointer");
  EXA_CHECK(hipSetDevice(g->s->device));
  EXA_CHECK(hipStreamSynchronize(g->stream));
  for (int p = 0; p < 3; ++p) {
    for (auto& e : g->ev[p]) {
      float ms = 0.f;
      EXA_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
      g->times[p] += ms;
      g->counts[p] += 1;
    }
  }
  grid_clear_events(g);
  for (int p = 0; p < 3; ++p) {
    times[p] = g->times[p];
    counts[p] = g->counts[p];
    g->times[p] = 0.0;
    g->counts[p] = 0;
  }
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// Ensemble of independent periodic grids (uneven sizes, orders and gammas):
// the instances of one order share every launch of a step -- one predictor,
// one Riemann and one update launch over all their cells, each cell reading
// its instance's dt, mesh width and gamma -- and the launches of all orders
// of one step form one hipGraph.
struct EnsGroup {
  int order = 0;
  exa_solver* s = nullptr;  // owned; carries the group's configuration
  long long nc = 0;
  std::vector<int> members;  // ensemble indices of the instances of this group
  std::vector<exa::InstParams> hp;
  DevBuf luh, lduh, qf, ff, iters, nbr, cell_inst, params, ics, iter_sum;
};

struct exa_ensemble {
  int device = 0;
  std::vector<EnsGroup*> groups;
  std::vector<std::pair<int, int>> where;  // instance -> (group, index in group)
  hipStream_t stream = nullptr;
  bool use_graph = true;
  bool dt_valid = false;  // exa_ensemble_compute_dt has run (sync checks the dts from then on)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

static exa::Ens ens_of(EnsGroup* gr) {
  exa::Ens e;
  e.cell_inst = gr->cell_inst.as<int>();
  e.inst = gr->params.as<exa::InstParams>();
  e.ics = gr->ics.as<exa_ic>();
  return e;
}

extern "C" int exa_ensemble_destroy(exa_ensemble* e) {
  if (!e) return EXA_OK;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  for (EnsGroup* gr : e->groups) {
    exa_aderdg_destroy(gr->s);
    delete gr;
  }
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return EXA_OK;
}

extern "C" int exa_ensemble_create(const exa_ensemble_member* m, int n, const exa_aderdg_config* cfg,
                                   exa_ensemble** out) {
  EXA_ARG(m && n > 0 && cfg && out, "arguments");
  EXA_ARG(cfg->contraction != EXA_CONTRACT_STAGED, "ensembles run the register-resident predictor forms");
  EXA_ARG(cfg->pde == EXA_PDE_EULER, "ensemble members solve the Euler equations without sources");
  size_t need = 0;
  for (int i = 0; i < n; ++i) {
    EXA_ARG(m[i].order >= 1 && m[i].order <= 7, "member order must be 1..7");
    EXA_ARG(m[i].nx > 0 && m[i].ny > 0 && m[i].nz > 0, "member grid dimensions");
    EXA_ARG(m[i].gamma > 1.0, "member gamma must exceed 1");
    const size_t B = m[i].order + 1, nc = (size_t)m[i].nx * m[i].ny * m[i].nz;
    need += nc * (2 * exa::V * B * B * B + 12 * exa::V * B * B) * sizeof(double) + nc * 5 * sizeof(int);
  }
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  size_t freeb = 0, total = 0;
  EXA_CHECK(hipMemGetInfo(&freeb, &total));
  if (need > freeb) {
    exa_set_error("ensemble needs %.1f GB of device memory, %.1f GB free: advance it in smaller waves",
                  need / 1e9, freeb / 1e9);
    return EXA_ERR_NOMEM;
  }
  exa_ensemble* e = new exa_ensemble();
  e->device = dev;
  e->where.resize(n);
  for (int N = 1; N <= 7; ++N) {
    std::vector<int> mem;
    for (int i = 0; i < n; ++i)
      if (m[i].order == N) mem.push_back(i);
    if (mem.empty()) continue;
    EnsGroup* gr = new EnsGroup();
    e->groups.push_back(gr);
    gr->order = N;
    gr->members = mem;
    exa_aderdg_config c = *cfg;
    c.order = N;
    c.nvar = 5;
    c.npar = 0;
    c.pde = EXA_PDE_EULER;
    if (c.cfl <= 0.0) c.cfl = 0.9;
    if (c.gamma <= 1.0) c.gamma = 1.4;
    int r = exa_aderdg_create(&c, &gr->s);
    if (r != EXA_OK) {
      exa_ensemble_destroy(e);
      return r;
    }
    long long base = 0;
    for (size_t j = 0; j < mem.size(); ++j) {
      const exa_ensemble_member& mm = m[mem[j]];
      exa::InstParams p{};
      for (int d = 0; d < 3; ++d) {
        p.off[d] = mm.offset[d];
        p.dx[d] = mm.width[d] / (d == 0 ? mm.nx : (d == 1 ? mm.ny : mm.nz));
      }
      p.gamma = mm.gamma;
      p.dt_cur = 0.0;
      p.dt_next = exa_big_bits();
      p.time = 0.0;
      p.base = base;
      p.nx = mm.nx;
      p.ny = mm.ny;
      p.nz = mm.nz;
      gr->hp.push_back(p);
      e->where[mem[j]] = {(int)e->groups.size() - 1, (int)j};
      base += (long long)mm.nx * mm.ny * mm.nz;
    }
    gr->nc = base;
    if (base >= (1LL << 31)) {
      exa_ensemble_destroy(e);
      exa_set_error("invalid argument: more than 2^31 cells of one order in an ensemble");
      return EXA_ERR_ARG;
    }
    const int B = N + 1, B2 = B * B, B3 = B2 * B, V = exa::V;
    if ((r = gr->luh.ensure(sizeof(double) * gr->nc * V * B3)) != EXA_OK ||
        (r = gr->lduh.ensure(sizeof(double) * gr->nc * V * B3)) != EXA_OK ||
        (r = gr->qf.ensure(sizeof(double) * gr->nc * 6 * V * B2)) != EXA_OK ||
        (r = gr->ff.ensure(sizeof(double) * gr->nc * 6 * V * B2)) != EXA_OK ||
        (r = gr->iters.ensure(sizeof(int) * gr->nc)) != EXA_OK ||
        (r = gr->nbr.ensure(sizeof(int) * 6 * gr->nc)) != EXA_OK ||
        (r = gr->cell_inst.ensure(sizeof(int) * gr->nc)) != EXA_OK ||
        (r = gr->params.ensure(sizeof(exa::InstParams) * mem.size())) != EXA_OK ||
        (r = gr->ics.ensure(sizeof(exa_ic) * mem.size())) != EXA_OK ||
        (r = gr->iter_sum.ensure(sizeof(unsigned long long))) != EXA_OK) {
      exa_ensemble_destroy(e);
      return r;
    }
    std::vector<int> nbr(6 * gr->nc), ci(gr->nc);
    for (size_t j = 0; j < mem.size(); ++j) {
      const exa::InstParams& p = gr->hp[j];
      const long long nci = (long long)p.nx * p.ny * p.nz;
      for (long long c = 0; c < nci; ++c) {
        const int ix = (int)(c % p.nx), iy = (int)((c / p.nx) % p.ny), iz = (int)(c / ((long long)p.nx * p.ny));
        const long long g = p.base + c;
        nbr[0 * gr->nc + g] = (int)(p.base + ((long long)iz * p.ny + iy) * p.nx + (ix + 1) % p.nx);
        nbr[1 * gr->nc + g] = (int)(p.base + ((long long)iz * p.ny + (iy + 1) % p.ny) * p.nx + ix);
        nbr[2 * gr->nc + g] = (int)(p.base + ((long long)((iz + 1) % p.nz) * p.ny + iy) * p.nx + ix);
        nbr[3 * gr->nc + g] = (int)(p.base + ((long long)iz * p.ny + iy) * p.nx + (ix + p.nx - 1) % p.nx);
        nbr[4 * gr->nc + g] = (int)(p.base + ((long long)iz * p.ny + (iy + p.ny - 1) % p.ny) * p.nx + ix);
        nbr[5 * gr->nc + g] = (int)(p.base + ((long long)((iz + p.nz - 1) % p.nz) * p.ny + iy) * p.nx + ix);
        ci[g] = (int)j;
      }
    }
    hipError_t he = hipMemcpy(gr->nbr.p, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(gr->cell_inst.p, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice);
    if (he == hipSuccess)
      he = hipMemcpy(gr->params.p, gr->hp.data(), sizeof(exa::InstParams) * gr->hp.size(), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemset(gr->iters.p, 0, sizeof(int) * gr->nc);
    if (he == hipSuccess) he = hipMemset(gr->iter_sum.p, 0, sizeof(unsigned long long));
    if (he != hipSuccess) {
      exa_set_error("ensemble setup: %s", hipGetErrorString(he));
      exa_ensemble_destroy(e);
      return EXA_ERR_HIP;
    }
  }
  hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    exa_set_error("hipStreamCreate: %s", hipGetErrorString(he));
    exa_ensemble_destroy(e);
    return EXA_ERR_HIP;
  }
  *out = e;
  return EXA_OK;
}

extern "C" int exa_ensemble_set_initial_conditions(exa_ensemble* e, const exa_ic* ics) {
  EXA_ARG(e && ics, "null pointer");
  EXA_CHECK(hipSetDevice(e->device));
  for (EnsGroup* gr : e->groups) {
    std::vector<exa_ic> h;
    for (int i : gr->members) h.push_back(ics[i]);
    EXA_CHECK(hipMemcpyAsync(gr->ics.p, h.data(), sizeof(exa_ic) * h.size(), hipMemcpyHostToDevice, e->stream));
    EXA_CHECK(hipStreamSynchronize(e->stream));
    const exa::Ens en = ens_of(gr);
    EXA_ORDER_SWITCH(gr->order, {
      using O = exa::Ord<NN>;
      hipLaunchKernelGGL((exa::k_initial_conditions<NN>), dim3((unsigned)gr->nc), dim3(O::NT), 0, e->stream,
                         gr->luh.as<double>(), ics[0], 1.4, 1, 1, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, en);
    });
    EXA_CHECK(hipGetLastError());
  }
  EXA_CHECK(hipStreamSynchronize(e->stream));
  return EXA_OK;
}

extern "C" int exa_ensemble_compute_dt(exa_ensemble* e) {
  EXA_ARG(e, "null pointer");
  EXA_CHECK(hipSetDevice(e->device));
  for (EnsGroup* gr : e->groups) {
    const int ni = (int)gr->members.size();
    const exa::Ens en = ens_of(gr);
    hipLaunchKernelGGL(exa::k_ens_dt_reset, dim3((ni + 255) / 256), dim3(256), 0, e->stream, en.inst, ni);
    EXA_ORDER_SWITCH(gr->order, {
      using O = exa::Ord<NN>;
      hipLaunchKernelGGL((exa::k_update<NN, exa::UPD_DT>), dim3((unsigned)gr->nc), dim3(O::NT), 0, e->stream,
                         gr->luh.as<double>(), gr->lduh.as<double>(), gr->ff.as<double>(), nullptr, nullptr, 1.0,
                         1.0, 1.0, 1.4, gr->s->cfg.cfl, en);
    });
    hipLaunchKernelGGL(exa::k_ens_dt_take, dim3((ni + 255) / 256), dim3(256), 0, e->stream, en.inst, ni);
    EXA_CHECK(hipGetLastError());
  }
  EXA_CHECK(hipStreamSynchronize(e->stream));
  e->dt_valid = true;
  return exa_ensemble_sync(e, nullptr, nullptr);
}

static int enqueue_ensemble_step(exa_ensemble* e) {
  exa::PredictorDebug nodbg;
  nodbg.lQi = nodbg.lFi = nodbg.lQhi = nodbg.lFhi = nullptr;
  for (EnsGroup* gr : e->groups) {
    exa::Ens en = ens_of(gr);
    en.fold_update = 1;  // k_riemann_update follows
    const exa_aderdg_config& cfg = gr->s->cfg;
    EXA_ORDER_SWITCH(gr->order, {
      using O = exa::Ord<NN>;
      const int wpe = predictor_waves_per_simd(cfg, NN, gr->s->device);
      const int form = predictor_form(cfg, NN, gr->s->device);
      PredFn kern = predictor_kernel<NN>(false, cfg.precision, wpe, form);
      hipLaunchKernelGGL(kern, dim3((unsigned)gr->nc), dim3(O::NT), O::LDS_DOUBLES * sizeof(double), e->stream,
                         gr->luh.as<double>(), gr->lduh.as<double>(), gr->qf.as<double>(), gr->ff.as<double>(),
                         gr->iters.as<int>(), nullptr, 1.0, 1.0, 1.0, 1.4, cfg.fp64_sweeps,
                         gr->iter_sum.as<unsigned long long>(), nodbg, en);
    });
    const double unit[3] = {1.0, 1.0, 1.0};  // per member from the instance table
    EXA_TRY(launch_riemann_update(gr->order, gr->luh.as<double>(), gr->lduh.as<double>(), gr->qf.as<double>(),
                                  gr->ff.as<double>(), gr->nbr.as<int>(), gr->nc, nullptr, nullptr, nullptr, unit, 1.4,
                                  cfg.cfl, e->stream, en));
    const int ni = (int)gr->members.size();
    hipLaunchKernelGGL(exa::k_ens_dt_advance, dim3((ni + 255) / 256), dim3(256), 0, e->stream, en.inst, ni);
    EXA_CHECK(hipGetLastError());
  }
  return EXA_OK;
}

extern "C" int exa_ensemble_use_graph(exa_ensemble* e, int enable) {
  EXA_ARG(e, "null pointer");
  e->use_graph = enable != 0;
  return EXA_OK;
}

extern "C" int exa_ensemble_step(exa_ensemble* e, int nsteps) {
  EXA_ARG(e && nsteps >= 0, "nsteps");
  EXA_CHECK(hipSetDevice(e->device));
  if (e->use_graph) {
    if (!e->gexec) {
      EXA_CHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      int r = enqueue_ensemble_step(e);
      hipGraph_t graph = nullptr;
      hipError_t he = hipStreamEndCapture(e->stream, &graph);
      if (r != EXA_OK) return r;
      EXA_CHECK(he);
      e->graph = graph;
      EXA_CHECK(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < nsteps; ++i) EXA_CHECK(hipGraphLaunch(e->gexec, e->stream));
  } else {
    for (int i = 0; i < nsteps; ++i) EXA_TRY(enqueue_ensemble_step(e));
  }
  EXA_CHECK(hipGetLastError());
  return EXA_OK;
}

extern "C" int exa_ensemble_sync(exa_ensemble* e, double* times, double* dts) {
  EXA_ARG(e, "null pointer");
  EXA_CHECK(hipSetDevice(e->device));
  EXA_CHECK(hipStreamSynchronize(e->stream));
  for (EnsGroup* gr : e->groups) {
    std::vector<exa::InstParams> p(gr->members.size());
    EXA_CHECK(hipMemcpy(p.data(), gr->params.p, sizeof(exa::InstParams) * p.size(), hipMemcpyDeviceToHost));
    for (size_t j = 0; j < p.size(); ++j) {
      if (times) times[gr->members[j]] = p[j].time;
      if (dts) dts[gr->members[j]] = p[j].dt_cur;
    }
    for (size_t j = 0; e->dt_valid && j < p.size(); ++j) {
      const int r = exa_check_dt(p[j].dt_cur, "exa_ensemble_sync");
      if (r != EXA_OK) {
        exa_set_error("exa_ensemble_sync: member %d has stable time step %g -- its state holds non-finite values "
                      "or a negative pressure", gr->members[j], p[j].dt_cur);
        return r;
      }
    }
  }
  return EXA_OK;
}

extern "C" int exa_ensemble_download(exa_ensemble* e, int instance, double* luh) {
  EXA_ARG(e && luh && instance >= 0 && instance < (int)e->where.size(), "instance");
  EXA_CHECK(hipSetDevice(e->device));
  EXA_CHECK(hipStreamSynchronize(e->stream));
  EnsGroup* gr = e->groups[e->where[instance].first];
  const exa::InstParams& p = gr->hp[e->where[instance].second];
  const int B = gr->order + 1, B3 = B * B * B, V = exa::V;
  const size_t per = (size_t)V * B3, nci = (size_t)p.nx * p.ny * p.nz;
  std::vector<double> buf(per * nci);
  EXA_CHECK(hipMemcpy(buf.data(), gr->luh.as<double>() + p.base * per, sizeof(double) * buf.size(),
                      hipMemcpyDeviceToHost));
  for (size_t c = 0; c < nci; ++c) soa_to_ref_cell(buf.data() + c * per, luh + c * per, B3, V);
  return EXA_OK;
}

extern "C" int exa_ensemble_total_picard_iterations(exa_ensemble* e, long long* total) {
  EXA_ARG(e && total, "null pointer");
  EXA_CHECK(hipSetDevice(e->device));
  EXA_CHECK(hipStreamSynchronize(e->stream));
  long long t = 0;
  for (EnsGroup* gr : e->groups) {
    unsigned long long h = 0, zero = 0;
    EXA_CHECK(hipMemcpy(&h, gr->iter_sum.p, sizeof(h), hipMemcpyDeviceToHost));
    EXA_CHECK(hipMemcpy(gr->iter_sum.p, &zero, sizeof(zero), hipMemcpyHostToDevice));
    t += (long long)h;
  }
  *total = t;
  return EXA_OK;
}

// ---------------------------------------------------------------------------
// Autotuner of the predictor's launch variant: times every (waves per SIMD,
// contraction form) pair of one order and precision on a synthetic vortex
// grid and records the fastest for this device.
extern "C" int exa_autotune(int order, int precision, int ncells, int* waves_per_simd, int* contraction,
                            double ms[6]) {
  EXA_ARG(order >= 1 && order <= 7, "order must be 1..7");
  EXA_ARG(precision >= EXA_PREC_FP64 && precision <= EXA_PREC_BF16_MFMA, "precision");
  EXA_ARG(ncells >= 4 && ncells <= 128, "ncells must be 4..128");
  int dev = 0;
  EXA_CHECK(hipGetDevice(&dev));
  EXA_TRY(exa_init(dev));
  exa_ic ic{};
  ic.kind = 0;
  ic.xc[0] = ic.xc[1] = ic.xc[2] = 0.5;
  ic.L = 0.1;
  ic.eps = 5.0;
  ic.uinf[0] = ic.uinf[1] = 1.0;
  ic.rho_inf = ic.p_inf = 1.0;
  double best = 1e300;
  int bw = 0, bc = 0, k = 0;
  for (int w = 1; w <= 2; ++w)
    for (int cform = EXA_CONTRACT_PENCIL; cform <= EXA_CONTRACT_PENCIL_PAIR; ++cform, ++k) {
      exa_aderdg_config cfg{};
      cfg.order = order;
      cfg.nvar = 5;
      cfg.cfl = 0.9;
      cfg.gamma = 1.4;
      cfg.pde = EXA_PDE_EULER;
      cfg.precision = precision;
      cfg.fp64_sweeps = 1;
      cfg.waves_per_simd = w;
      cfg.contraction = cform;
      exa_solver* s = nullptr;
      exa_grid* g = nullptr;
      EXA_TRY(exa_aderdg_create(&cfg, &s));
      int r = exa_grid_create(s, ncells, ncells, ncells, nullptr, nullptr, &g);
      double t = 1e300;
      if (r == EXA_OK) r = exa_grid_set_initial_conditions(g, &ic);
      if (r == EXA_OK) r = exa_grid_compute_dt(g, nullptr);
      if (r == EXA_OK) r = exa_grid_step(g, 1);
      if (r == EXA_OK) r = exa_grid_enable_timing(g, 1);
      if (r == EXA_OK) {
        double tt[3];
        long long cc[3];
        r = exa_grid_kernel_times(g, tt, cc);
        if (r == EXA_OK) r = exa_grid_step(g, 3);
        if (r == EXA_OK) r = exa_grid_kernel_times(g, tt, cc);
        if (r == EXA_OK && cc[0] > 0) t = tt[0] / cc[0];
      }
      exa_grid_destroy(g);
      exa_aderdg_destroy(s);
      if (r != EXA_OK) return r;
      if (ms) ms[k] = t;
      if (t < best) {
        best = t;
        bw = w;
        bc = cform;
      }
    }
  g_tuned[dev & 63][order][precision].wpe = bw;
  g_tuned[dev & 63][order][precision].contraction = bc;
  if (waves_per_simd) *waves_per_simd = bw;
  if (contraction) *contraction = bc;
  return EXA_OK;
}
